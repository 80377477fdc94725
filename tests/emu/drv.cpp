// Driver for the CPU emulation of k_bsr_subtree and k_radix_subtree; the
// kernels' source is cut out of bluecrunch_amd/csrc/e_pipeline.inc by
// tests/test_subtree_emulation.py into kernels_cpu.h.
//   drv bsr T out        P, Q of every subtree of BSR_SUB leaves for T terms
//   drv radix in out     digit groups of the roots described in `in`
#include <algorithm>
#include <fstream>
#include <string>

#include "kernels_cpu.h"

int main(int argc, char **argv) {
  if (argc < 4) return 2;
  std::string mode = argv[1];
  if (mode == "bsr") {
    unsigned long long T = std::stoull(argv[2]);
    unsigned long long leaves = (T + BSR_G - 1) / BSR_G, nodes = (leaves + BSR_SUB - 1) / BSR_SUB;
    std::vector<uint32_t> buf(nodes * 2 * BSR_SUB_STRIDE, 0xdeadbeefu);
    std::vector<unsigned long long> len(2 * nodes);
    run_grid((unsigned)nodes, BSR_SUB_T, [&] { k_bsr_subtree(T, leaves, buf.data(), len.data()); });
    std::ofstream o(argv[3]);
    for (unsigned long long i = 0; i < nodes; i++)
      for (int v = 0; v < 2; v++) {
        o << len[2 * i + v];
        for (unsigned long long t = 0; t < len[2 * i + v]; t++) o << ' ' << buf[(2 * i + v) * BSR_SUB_STRIDE + t];
        o << '\n';
      }
    return 0;
  }
  // input: m, then per root "K n d0 words..."; then npow, per power "K len words..."
  std::ifstream in(argv[2]);
  size_t m;
  in >> m;
  std::vector<unsigned long long> off(m), n(m), K(m), d0(m);
  std::vector<uint32_t> buf;
  size_t D = 0;
  for (size_t i = 0; i < m; i++) {
    in >> K[i] >> n[i] >> d0[i];
    off[i] = buf.size();
    for (size_t t = 0; t < n[i]; t++) {
      unsigned long long w;
      in >> w;
      buf.push_back((uint32_t)w);
    }
    D = std::max<size_t>(D, d0[i] + K[i]);
  }
  int np;
  in >> np;
  std::vector<RPow> tab(np);
  std::vector<uint32_t> pw;
  for (int i = 0; i < np; i++) {
    unsigned k, l;
    in >> k >> l;
    tab[i] = {k, (unsigned)pw.size(), l};
    for (unsigned t = 0; t < l; t++) {
      unsigned long long w;
      in >> w;
      pw.push_back((uint32_t)w);
    }
  }
  std::vector<uint32_t> dig(D, 0xffffffffu);
  run_grid((unsigned)m, RADIX_SUB_T, [&] {
    k_radix_subtree(buf.data(), off.data(), n.data(), K.data(), d0.data(), pw.data(), tab.data(), np, dig.data());
  });
  std::ofstream o(argv[3]);
  for (size_t i = 0; i < D; i++) o << dig[i] << '\n';
  return 0;
}
