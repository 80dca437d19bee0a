// swm_bf16.hip -- the step and ensemble launches for the number-format
// bundle FmtBF16, in a translation unit of their own (the units compile in parallel).
#include "swm_launch.h"

const FmtOps &ops_bf16() {
    static const FmtOps o = Entry<FmtBF16, false>::make();
    return o;
}
