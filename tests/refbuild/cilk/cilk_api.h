/* Compile-only stand-in for the two Cilk Plus runtime calls the reference's
 * callers make (worker count query, worker count setting); see cilk.h here. */
#ifndef REFBUILD_CILK_API_H
#define REFBUILD_CILK_API_H
static inline int __cilkrts_get_nworkers(void) { return 1; }
static inline int __cilkrts_set_param(const char *name, const char *value) {
  (void)name;
  (void)value;
  return 0;
}
#endif
