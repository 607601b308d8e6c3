// Stand-in for rocPRIM's segmented radix sorts in the CPU emulation of
// tests/test_select_emulation.py: a stable sort of each segment, with the
// segment bounds checked against the buffer size.
#pragma once
#include <hip/hip_runtime.h>
#include <numeric>
namespace rocprim {
template <class KI, class KO, class OI>
hipError_t segmented_radix_sort_keys(void *tmp, size_t &bytes, KI in, KO out, unsigned size, unsigned segs, OI b, OI e,
                                     unsigned, unsigned, hipStream_t) {
  if (!tmp) { bytes = 64; return 0; }
  for (unsigned i = 0; i < segs; ++i) {
    if (!(b[i] <= e[i] && (unsigned long long)e[i] <= size)) { fprintf(stderr, "EMU sort keys: bad segment %u [%lld,%lld) size %u\n", i, (long long)b[i], (long long)e[i], size); abort(); }
    for (auto j = b[i]; j < e[i]; ++j) out[j] = in[j];
    std::stable_sort(out + b[i], out + e[i]);
  }
  return 0;
}
template <class KI, class KO, class VI, class VO, class OI>
hipError_t segmented_radix_sort_pairs(void *tmp, size_t &bytes, KI kin, KO kout, VI vin, VO vout, unsigned size,
                                      unsigned segs, OI b, OI e, unsigned, unsigned, hipStream_t) {
  if (!tmp) { bytes = 64; return 0; }
  for (unsigned i = 0; i < segs; ++i) {
    if (!(b[i] <= e[i] && (unsigned long long)e[i] <= size)) { fprintf(stderr, "EMU sort pairs: bad segment\n"); abort(); }
    std::vector<long long> idx(e[i] - b[i]);
    std::iota(idx.begin(), idx.end(), (long long)b[i]);
    std::stable_sort(idx.begin(), idx.end(), [&](long long x, long long y) { return kin[x] < kin[y]; });
    std::vector<decltype(+kin[0])> ks;
    std::vector<decltype(+vin[0])> vs;
    for (auto j : idx) { ks.push_back(kin[j]); vs.push_back(vin[j]); }
    for (size_t j = 0; j < idx.size(); ++j) { kout[b[i] + j] = ks[j]; vout[b[i] + j] = vs[j]; }
  }
  return 0;
}
}  // namespace rocprim
