// One schedule variant of the G1 transform kernels, chosen by the build: FK20_VW (wNAF width
// of the twiddle programs, 3..5), FK20_VR22 (1: radix-2^2 butterflies and twiddle layers,
// 0: radix 2), FK20_VOCC (waves per SIMD the register allocation allows).  The Makefile
// compiles this file once per variant; fast_kernels.hip looks the tables up by name.
#include "stage_kernels.hip.h"

#define FK20_PASTE(a, w, r, o) a##w##_##r##_##o
#define FK20_VARIANT(w, r, o) FK20_PASTE(stage_kernels_, w, r, o)

namespace fk20 {
extern const StageKernels FK20_VARIANT(FK20_VW, FK20_VR22, FK20_VOCC);
const StageKernels FK20_VARIANT(FK20_VW, FK20_VR22, FK20_VOCC) =
    make_stage_kernels<FK20_VW, FK20_VR22 != 0, FK20_VOCC>();
}  // namespace fk20
