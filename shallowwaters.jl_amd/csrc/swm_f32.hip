// swm_f32.hip -- the step, ensemble and reverse-mode launches for the number-format
// bundle FmtF32, in a translation unit of their own (the units compile in parallel).
#include "swm_launch.h"

const FmtOps &ops_f32() {
    static const FmtOps o = Entry<FmtF32, true>::make();
    return o;
}
