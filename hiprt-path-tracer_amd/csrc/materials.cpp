// materials.cpp -- CPUMaterial -> packed GPU material -> unpacked floats.
//
// The reference uploads every material in an 8-bit packed form
// (CPUMaterial::pack_to_GPU, src/HostDeviceCommon/Material/MaterialCPU.h:31-150;
// packers in src/HostDeviceCommon/Packing.h:56-203) and every kernel reads
// the unpacked values. The round trip is done once here, on the host, with
// the same binary32 arithmetic: truncation u8(x * 255) on the way in;
// u8 * (1/255) on the way out, except the green channel of a 3x8-bit colour,
// which the reference unpacks as float(g << 8) * (1 / (255 << 8)).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/mipt.h"

namespace mipt {
namespace {

const float INV_255 = 1.0f / 255.0f;
const float INV_255_SHL_8 = 1.0f / (float)(255 << 8);

uint8_t q8(float x) { return static_cast<unsigned char>(x * 255.0f); }
float u8f(uint8_t u) { return static_cast<float>(u) * INV_255; }
void color3(const float in[3], float out[3])
{
    uint32_t r = q8(in[0]), g = q8(in[1]), b = q8(in[2]);
    out[0] = static_cast<float>(r) * INV_255;
    out[1] = static_cast<float>(g << 8) * INV_255_SHL_8;
    out[2] = static_cast<float>(b) * INV_255;
}

float clampf(float lo, float hi, float v) { return std::min(hi, std::max(lo, v)); }
void clamp3(float c[3], float lo, float hi)
{
    for (int k = 0; k < 3; k++)
        c[k] = clampf(lo, hi, c[k]);
}

// CPUMaterial::make_safe (src/HostDeviceCommon/Material/MaterialCPU.h:156-212),
// which the scene parser applies to every material (SceneParser.cpp:437).
mipt_material make_safe(mipt_material m)
{
    clamp3(m.base_color, 0.0f, 1.0f);
    m.metallic = clampf(0.0f, 1.0f, m.metallic);
    clamp3(m.metallic_F82, 0.0f, 1.0f);
    clamp3(m.metallic_F90, 0.0f, 1.0f);
    m.anisotropy = clampf(0.0f, 1.0f, m.anisotropy);
    m.anisotropy_rotation = clampf(0.0f, 1.0f, m.anisotropy_rotation);
    m.second_roughness_weight = clampf(0.0f, 1.0f, m.second_roughness_weight);
    m.second_roughness = clampf(0.0f, 1.0f, m.second_roughness);
    m.specular = clampf(0.0f, 1.0f, m.specular);
    m.specular_tint = clampf(0.0f, 1.0f, m.specular_tint);
    clamp3(m.specular_color, 0.0f, 1.0f);
    m.specular_darkening = clampf(0.0f, 1.0f, m.specular_darkening);
    m.coat = clampf(0.0f, 1.0f, m.coat);
    clamp3(m.coat_medium_absorption, 0.0f, 1.0f);
    m.coat_roughness = clampf(1.0e-4f, 1.0f, m.coat_roughness);
    m.coat_roughening = clampf(0.0f, 1.0f, m.coat_roughening);
    m.coat_darkening = clampf(0.0f, 1.0f, m.coat_darkening);
    m.coat_anisotropy = clampf(0.0f, 1.0f, m.coat_anisotropy);
    m.coat_anisotropy_rotation = clampf(0.0f, 1.0f, m.coat_anisotropy_rotation);
    m.sheen = clampf(0.0f, 1.0f, m.sheen);
    m.sheen_roughness = clampf(1.0e-4f, 1.0f, m.sheen_roughness);
    clamp3(m.sheen_color, 0.0f, 1.0f);
    m.specular_transmission = clampf(0.0f, 1.0f, m.specular_transmission);
    m.absorption_at_distance = std::max(m.absorption_at_distance, 1.0e-4f);
    for (int k = 0; k < 3; k++)
        m.absorption_color[k] = std::max(m.absorption_color[k], 1.0f / 512.0f);
    clamp3(m.absorption_color, 0.0f, 1.0f);
    m.dispersion_abbe_number = std::max(1.0e-5f, m.dispersion_abbe_number);
    m.dispersion_scale = clampf(0.0f, 1.0f, m.dispersion_scale);
    m.thin_film = clampf(0.0f, 1.0f, m.thin_film);
    m.thin_film_hue_shift_degrees = clampf(0.0f, 360.0f, m.thin_film_hue_shift_degrees);
    m.thin_film_ior = std::max(1.0005f, m.thin_film_ior);
    m.alpha_opacity = clampf(0.0f, 1.0f, m.alpha_opacity);
    m.dielectric_priority = std::min(15, std::max(0, m.dielectric_priority));
    for (int k = 0; k < 3; k++)
        m.emission[k] = std::max(0.0f, m.emission[k]);
    // an opaque boundary is never skipped: maximum priority, (1 << 15) - 1,
    // which the packed material keeps in 8 bits (MaterialPacked.h)
    if (m.specular_transmission == 0.0f && m.diffuse_transmission == 0.0f)
        m.dielectric_priority = (1 << 15) - 1;
    return m;
}

} // namespace

// Output layout = pt::Material (csrc/shading.h), 64 floats.
std::string unpack_material(const mipt_material& in, float o[64])
{
    const mipt_material m = make_safe(in);
    std::memset(o, 0, 64 * sizeof(float));
    // emission is stored at full precision, pre-multiplied by its strength
    // and the global factor (MaterialCPU.h:64)
    for (int k = 0; k < 3; k++)
        o[k] = m.emission[k] * m.emission_strength * m.global_emissive_factor;
    color3(m.base_color, o + 3);
    o[6] = u8f(q8(m.roughness));
    o[7] = u8f(q8(m.metallic));
    color3(m.metallic_F82, o + 8);
    color3(m.metallic_F90, o + 11);
    o[14] = m.metallic_F90_falloff_exponent;
    o[15] = u8f(q8(m.anisotropy));
    o[16] = u8f(q8(m.anisotropy_rotation));
    o[17] = u8f(q8(m.second_roughness_weight));
    o[18] = u8f(q8(m.second_roughness));
    o[19] = u8f(q8(m.specular));
    o[20] = u8f(q8(m.coat));
    o[21] = u8f(q8(m.sheen));
    o[22] = u8f(q8(m.specular_transmission));
    o[23] = u8f(q8(m.diffuse_transmission));
    o[24] = u8f(q8(m.thin_film));
    o[25] = m.ior;
    uint32_t flags = (m.do_metallic_energy_compensation ? 1u : 0u) | (m.do_specular_energy_compensation ? 2u : 0u) |
                     (m.do_coat_energy_compensation ? 4u : 0u) | (m.do_glass_energy_compensation ? 8u : 0u) |
                     (m.thin_walled ? 16u : 0u) | (m.thin_film_do_ior_override ? 32u : 0u);
    std::memcpy(&o[26], &flags, 4);
    o[27] = u8f(q8(m.specular_tint));
    color3(m.specular_color, o + 28);
    o[31] = u8f(q8(m.specular_darkening));
    color3(m.coat_medium_absorption, o + 32);
    o[35] = m.coat_medium_thickness;
    o[36] = u8f(q8(m.coat_roughness));
    o[37] = u8f(q8(m.coat_roughening));
    o[38] = u8f(q8(m.coat_darkening));
    o[39] = u8f(q8(m.coat_anisotropy));
    o[40] = u8f(q8(m.coat_anisotropy_rotation));
    o[41] = m.coat_ior;
    o[42] = u8f(q8(m.sheen_roughness));
    color3(m.sheen_color, o + 43);
    o[46] = m.absorption_at_distance;
    color3(m.absorption_color, o + 47);
    o[50] = u8f(q8(m.dispersion_scale));
    o[51] = m.dispersion_abbe_number;
    const int32_t priority = (uint8_t)m.dielectric_priority;
    std::memcpy(&o[52], &priority, 4);
    // thin film: IOR, thickness, kappa and override IOR at full precision, the
    // hue shift as 8 bits of degrees / 360 (MaterialCPU.h:125-132)
    o[53] = m.thin_film_ior;
    o[54] = m.thin_film_thickness;
    o[55] = m.thin_film_kappa_3;
    o[56] = u8f(q8(m.thin_film_hue_shift_degrees / 360.0f));
    o[57] = m.thin_film_base_ior_override;

    // Coat roughening of the base and second metal roughness
    // (get_intersection_material, src/Device/includes/Material.h:122-149).
    const float coat = o[20], coat_roughening = o[37], coat_roughness = o[36];
    if (coat > 0.0f && coat_roughening > 0.0f)
    {
        auto pow4 = [](float x) { float x2 = x * x; return x2 * x2; };
        auto lerp = [](float a, float b, float t) { return (1.0f - t) * a + t * b; };
        const float base = o[6];
        float target = sqrtf(sqrtf(std::min(1.0f, pow4(base) + 2.0f * pow4(coat_roughness))));
        float roughened = lerp(base, target, coat);
        o[6] = lerp(base, roughened, coat_roughening);
        if (o[17] > 0.0f)
        {
            const float second = o[18];
            float target2 = sqrtf(sqrtf(std::min(1.0f, pow4(second) + 2.0f * pow4(coat_roughness))));
            float roughened2 = lerp(second, target2, coat);
            o[18] = lerp(second, roughened2, coat_roughening);
        }
    }

    return "";
}

} // namespace mipt
