// Kernel variants: every pass kind is built for runtime lengths (size 0)
// and for fixed lengths, for which the FFT stages unroll into straight-line
// code with every index a constant (about twice as fast as the runtime
// engine).  The fixed lengths are the 2^a 3^b 5^c sizes of the batched and
// served workloads:
//
//   KW_FIXED_SIZES_ALL: every pass kind (any axis of a 2D or 3D grid),
//                       64 .. 256 and 384;
//   KW_FIXED_SIZES_XL:  every kind except the strided y passes (the x axis
//                       and the last axis of any grid, which covers 2D
//                       grids), 270 .. 1024.
//
// The host asks the backend (fixed_variant in kw_backend.h) which variant
// exists for a pass and an axis length.
#pragma once

#define KW_FIXED_SIZES_ALL(X)                                                                    \
  X(64) X(72) X(75) X(80) X(81) X(90) X(96) X(100) X(108) X(120) X(125) X(128) X(135) X(144) \
  X(150) X(160) X(162) X(180) X(192) X(200) X(216) X(225) X(240) X(243) X(250) X(256) X(384)

#define KW_FIXED_SIZES_XL(X)                                                                     \
  X(270) X(288) X(300) X(320) X(324) X(360) X(375) X(400) X(405) X(432) X(450) X(480) X(486) \
  X(500) X(512) X(540) X(576) X(600) X(625) X(640) X(648) X(675) X(720) X(729) X(750) X(768) \
  X(800) X(810) X(864) X(900) X(960) X(972) X(1000) X(1024)

// the emulator build instantiates a representative subset (one translation
// unit holds every variant it has): powers of two, lengths that are not
// multiples of the 8 elements per thread, odd lengths, and both lists
#define KW_EMU_FIXED_SIZES(X) X(64) X(75) X(81) X(128) X(135) X(256) X(270) X(384)

// launcher of one (pass kind, size) variant; dim selects the 2D/3D body
#define KW_DECLARE_VARIANT(KIND, SIZE)                                              \
  void launch_v_##KIND##_##SIZE(int dim, int nb, int T, size_t lds, void* s,      \
                                const Batch& B, const PassArgs& A);
