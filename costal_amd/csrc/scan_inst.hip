// scan_inst.hip -- instantiates the fused scan kernels of one precision form and list capacity.
// Built once per (SCAN_PREC, SCAN_KP) pair (see the Makefile), so the nine kernel families compile
// in parallel; the host code picks one through the scan_fn_p<P>_k<K> selectors declared in scan.h.
// With SCAN_TANI, the Tanimoto-selection scan of split FP16 rows instead (scan_tani_fn_k<K>).
#include "scan.h"

#include <type_traits>

#ifndef SCAN_PREC
#error "SCAN_PREC must be defined (0 fp32, 1 fp16, 2 split fp16)"
#endif
#ifndef SCAN_KP
#error "SCAN_KP must be defined (16, 32 or 64)"
#endif

namespace {

#ifdef SCAN_TANI
static_assert(SCAN_PREC == PREC_F16X3, "Tanimoto scans run on split FP16 rows");
constexpr int SCAN_SEL = SEL_TANI;
#else
constexpr int SCAN_SEL = SEL_COS;
#endif

template <int J>
costal::CoarseFn one() {
  if constexpr (SCAN_PREC == PREC_F16X3 && J % 2) return nullptr;  // split rows hold hi and lo halves
  else return coarse_kernel<SCAN_PREC, J, SCAN_KP, 4, SCAN_SEL>;
}

template <int J>
costal::CoarseBatchFn one_batch() {
  if constexpr (SCAN_PREC == PREC_F16X3 && J % 2) return nullptr;
  else return coarse_batch_kernel<SCAN_PREC, J, SCAN_KP, 4>;
}

template <class F, class G>
F by_j(uint32_t J, G pick) {
  switch (J) {
    case 1: return pick(std::integral_constant<int, 1>{});
    case 2: return pick(std::integral_constant<int, 2>{});
    case 3: return pick(std::integral_constant<int, 3>{});
    case 4: return pick(std::integral_constant<int, 4>{});
    case 5: return pick(std::integral_constant<int, 5>{});
    case 6: return pick(std::integral_constant<int, 6>{});
    case 7: return pick(std::integral_constant<int, 7>{});
    case 8: return pick(std::integral_constant<int, 8>{});
  }
  return nullptr;
}

}  // namespace

#define COSTAL_CAT3(a, b, c, d) a##b##c##d
#define COSTAL_NAME(a, b, c, d) COSTAL_CAT3(a, b, c, d)

namespace costal {

#ifdef SCAN_TANI
CoarseFn COSTAL_NAME(scan_tani_fn, _k, SCAN_KP, )(uint32_t J) {
  return by_j<CoarseFn>(J, [](auto j) { return one<decltype(j)::value>(); });
}
#else
CoarseFn COSTAL_NAME(scan_fn_p, SCAN_PREC, _k, SCAN_KP)(uint32_t J) {
  return by_j<CoarseFn>(J, [](auto j) { return one<decltype(j)::value>(); });
}

CoarseBatchFn COSTAL_NAME(scan_batch_fn_p, SCAN_PREC, _k, SCAN_KP)(uint32_t J) {
  return by_j<CoarseBatchFn>(J, [](auto j) { return one_batch<decltype(j)::value>(); });
}
#endif

}  // namespace costal
