// Kernel variants: every pass kind is built for runtime lengths (size 0)
// and for each fixed length below, for which the FFT stages unroll into
// straight-line code.  The host picks the fixed variant when the axis length
// of the pass matches.
#pragma once

#define KW_FIXED_SIZES(X) X(64) X(128) X(256) X(384)

#define KW_IS_FIXED_SIZE(n) ((n) == 64 || (n) == 128 || (n) == 256 || (n) == 384)

// launcher of one (pass kind, size) variant; dim selects the 2D/3D body
#define KW_DECLARE_VARIANT(KIND, SIZE)                                              \
  void launch_v_##KIND##_##SIZE(int dim, int nb, int T, size_t lds, void* s,      \
                                const Batch& B, const PassArgs& A);
