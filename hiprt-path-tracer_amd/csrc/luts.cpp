// luts.cpp -- the BSDF tables embedded in the library, converted from the
// reference's baked data by data/convert_luts.py:
//   - GGX conductor directional albedo E_ss(cos_theta, roughness), 128 x 128
//     (and its height-correlated twin, which only mipt_get_shipped_lut reads);
//   - glossy dielectric (specular over diffuse) directional albedo, 128 layers
//     (F0^(1/4)) of 64 (roughness) x 128 (cos_theta^(1/2.5));
//   - GGX glass directional albedo and its inverse-IOR twin, 128 layers of
//     16 x 256, and the thin-walled glass albedo, 96 layers of 32 x 32;
//   - the glass correction exponent, a piecewise-linear function of the
//     roughness tabulated at ten relative IORs, [2][10][11][4];
//   - fitted sheen LTC parameters (Ai, Bi, Ri), 32 x 32 float3.
// All binary32, height-uncorrelated Smith masking (the reference's default).
#define MIPT_INCBIN(sym, file)                                                                                   \
    __asm__(".section .rodata\n"                                                                                 \
            ".balign 64\n"                                                                                       \
            ".global " #sym "\n" #sym ":\n"                                                                      \
            ".incbin \"" MIPT_DATA_DIR "/" file "\"\n"                                                           \
            ".previous\n")

MIPT_INCBIN(mipt_lut_ggx_conductor_uncorrelated, "ggx_conductor_uncorrelated_128x128.f32");
MIPT_INCBIN(mipt_lut_ggx_conductor_correlated, "ggx_conductor_correlated_128x128.f32");
MIPT_INCBIN(mipt_lut_glossy_dielectric_uncorrelated, "glossy_dielectric_uncorrelated_128x64x128.f32");
MIPT_INCBIN(mipt_lut_ggx_glass_uncorrelated, "ggx_glass_uncorrelated_256x16x128.f32");
MIPT_INCBIN(mipt_lut_ggx_glass_inv_uncorrelated, "ggx_glass_inv_uncorrelated_256x16x128.f32");
MIPT_INCBIN(mipt_lut_ggx_thin_glass_uncorrelated, "ggx_thin_glass_uncorrelated_32x32x96.f32");
MIPT_INCBIN(mipt_lut_glass_correction_exponent, "glass_correction_exponent_2x10x11x4.f32");
MIPT_INCBIN(mipt_lut_sheen_ltc, "sheen_ltc_32x32.f32");

extern "C" const float mipt_lut_ggx_conductor_uncorrelated[128 * 128];
extern "C" const float mipt_lut_ggx_conductor_correlated[128 * 128];
extern "C" const float mipt_lut_glossy_dielectric_uncorrelated[128 * 64 * 128];
extern "C" const float mipt_lut_ggx_glass_uncorrelated[256 * 16 * 128];
extern "C" const float mipt_lut_ggx_glass_inv_uncorrelated[256 * 16 * 128];
extern "C" const float mipt_lut_ggx_thin_glass_uncorrelated[32 * 32 * 96];
extern "C" const float mipt_lut_glass_correction_exponent[2 * 10 * 11 * 4];
extern "C" const float mipt_lut_sheen_ltc[32 * 32 * 3];

namespace mipt {
const float* conductor_lut_data() { return mipt_lut_ggx_conductor_uncorrelated; }
const float* conductor_correlated_lut_data() { return mipt_lut_ggx_conductor_correlated; }
const float* glossy_dielectric_lut_data() { return mipt_lut_glossy_dielectric_uncorrelated; }
const float* glass_lut_data() { return mipt_lut_ggx_glass_uncorrelated; }
const float* glass_inv_lut_data() { return mipt_lut_ggx_glass_inv_uncorrelated; }
const float* thin_glass_lut_data() { return mipt_lut_ggx_thin_glass_uncorrelated; }
const float* glass_exponent_data() { return mipt_lut_glass_correction_exponent; }
const float* sheen_ltc_data() { return mipt_lut_sheen_ltc; }
} // namespace mipt
