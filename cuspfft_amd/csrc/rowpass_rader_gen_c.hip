// rowpass_rader_gen_c.hip -- k_rowrader (FP32) for the generated shapes of list C
// (rader_shapes.h, written by gen_rader_shapes.py): odd prime widths p and widths 2p
// whose p - 1 is 7-smooth.
#include "rader_impl.h"
#include "rader_shapes.h"

namespace spfft {

#define SPFFT_GEN_LIST SPFFT_RADER_GEN_LIST_C
#define SPFFT_GEN_FN(name) name##_c
#include "rader_gen_impl.h"
#undef SPFFT_GEN_FN
#undef SPFFT_GEN_LIST

}  // namespace spfft
