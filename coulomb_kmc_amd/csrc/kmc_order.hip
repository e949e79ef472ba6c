// The launchers of the kernels templated on the expansion order, for one order:
// compiled once per L in [2, 20] with -DKMC_ORDER=L (kmc_order_fns_<L> is looked up
// by order_fns in kmc_amd.hip).
#include "kmc_device.h"

#ifndef KMC_ORDER
#error "build with -DKMC_ORDER=<expansion order>"
#endif
#define KMC_PASTE2(a, b) a##b
#define KMC_PASTE(a, b) KMC_PASTE2(a, b)

OrderFns KMC_PASTE(kmc_order_fns_, KMC_ORDER)() { return fns_for<KMC_ORDER>(); }
