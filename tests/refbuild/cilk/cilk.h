/* Compile-only stand-in for the Cilk Plus keywords, used by
 * tests/test_reference_link.py to compile the reference's callers of the
 * library (the Cilk Plus runtime is not installed here). This is the serial
 * elision that defines Cilk Plus's semantics: spawn and sync vanish, cilk_for
 * is a for loop. The executables it produces are linked, never run. */
#ifndef REFBUILD_CILK_H
#define REFBUILD_CILK_H
#define cilk_spawn
#define cilk_sync
#define cilk_for for
#endif
