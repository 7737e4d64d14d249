// luts.cpp -- the GGX conductor directional-albedo LUT (E_ss(cos_theta,
// roughness), 128 x 128 binary32, height-uncorrelated Smith masking, the
// reference's default flavour) embedded in the library. Source data:
// hiprt-path-tracer_amd/data/ggx_conductor_uncorrelated_128x128.f32, converted
// from the reference's baked .hdr by data/convert_luts.py.
__asm__(".section .rodata\n"
        ".balign 64\n"
        ".global mipt_lut_ggx_conductor_uncorrelated\n"
        "mipt_lut_ggx_conductor_uncorrelated:\n"
        ".incbin \"" MIPT_DATA_DIR "/ggx_conductor_uncorrelated_128x128.f32\"\n"
        ".previous\n");

extern "C" const float mipt_lut_ggx_conductor_uncorrelated[128 * 128];

namespace mipt {
const float* conductor_lut_data() { return mipt_lut_ggx_conductor_uncorrelated; }
} // namespace mipt
