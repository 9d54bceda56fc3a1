// context.h -- library-internal queries on a bnsamd_context (capi.hip) for
// the other translation units of the library.
#ifndef BNS_NURATES_AMD_CONTEXT_H_
#define BNS_NURATES_AMD_CONTEXT_H_

#include "../../include/bns_nurates_amd.h"

namespace bns {
// orders of the context's current 1D and 2D quadratures; false when none is set
bool context_quadrature_orders(bnsamd_context* ctx, int* n1, int* n2);
}  // namespace bns

#endif  // BNS_NURATES_AMD_CONTEXT_H_
