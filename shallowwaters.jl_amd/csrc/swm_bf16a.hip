// swm_bf16a.hip -- the step and ensemble launches for the number-format bundle FmtBF16A
// (BFloat16 values with float arithmetic, Tprog = Float32), in a translation unit of their own.
#include "swm_launch.h"

const FmtOps &ops_bf16a() {
    static const FmtOps o = Entry<FmtBF16A, false>::make();
    return o;
}
