// volume.h -- per-path volume state: the nested-dielectrics interior stack
// (priority-based nested dielectrics of Schmidt & Budge 2002), the distance
// travelled inside the current medium (Beer-Lambert absorption) and the
// wavelength sampled for dispersion.
//
// Behaviour follows NestedDielectricsInteriorStack
// (src/Device/includes/NestedDielectrics.h:125-255) and RayVolumeState
// (src/Device/includes/RayVolumeState.h:14-80) with a stack of 4 entries
// (NESTED_DIELECTRICS_STACK_SIZE, src/HostDeviceCommon/KernelOptions/KernelOptions.h:48).
// Entry 0 is the air and never changes. Every surface hit pushes its material;
// a reflection pops it again, a refraction out of a medium pops both entries.
#pragma once

#include "pt_math.h"

namespace pt {

constexpr int32_t VOL_STACK_SIZE = 4;
// Entry bits: priority 0-3, topmost 4, odd parity 5, material index 6-31.
constexpr uint32_t VOL_PRIO_MASK = 0xfu;
constexpr uint32_t VOL_TOPMOST = 1u << 4;
constexpr uint32_t VOL_ODD = 1u << 5;
constexpr uint32_t VOL_MAT_SHIFT = 6;
constexpr uint32_t VOL_AIR = 0xffffffffu >> VOL_MAT_SHIFT; // material index of the air

PT_HD uint32_t vol_mat(uint32_t e) { return e >> VOL_MAT_SHIFT; }
PT_HD uint32_t vol_prio(uint32_t e) { return e & VOL_PRIO_MASK; }
PT_HD bool vol_odd(uint32_t e) { return (e & VOL_ODD) != 0u; }
PT_HD bool vol_topmost(uint32_t e) { return (e & VOL_TOPMOST) != 0u; }
PT_HD uint32_t vol_entry(uint32_t mat, uint32_t prio, bool odd, bool topmost)
{
    return ((mat & VOL_AIR) << VOL_MAT_SHIFT) | (odd ? VOL_ODD : 0u) | (topmost ? VOL_TOPMOST : 0u) |
           (prio & VOL_PRIO_MASK);
}

// The four entries are separate fields read and written through selects: a
// dynamically indexed array would be placed in scratch memory.
struct VolumeState {
    uint32_t e0, e1, e2, e3;
    int32_t pos;       // last valid entry
    int32_t incident;  // medium the ray arrives from: material index or VOL_AIR (-1 before any push)
    int32_t outgoing;  // medium on the other side of the surface
    int32_t inside;    // the hit leaves the medium of its material
    float distance;    // distance travelled in the current medium
    float wavelength;  // 0: none; < 0: sampled, throughput filter not applied yet
    int32_t reset;     // set when a refraction or transmission reset the distance (not stored)
};

PT_HD uint32_t vol_get(const VolumeState& v, int i) { return i == 0 ? v.e0 : i == 1 ? v.e1 : i == 2 ? v.e2 : v.e3; }
PT_HD void vol_set(VolumeState& v, int i, uint32_t x)
{
    v.e0 = i == 0 ? x : v.e0;
    v.e1 = i == 1 ? x : v.e1;
    v.e2 = i == 2 ? x : v.e2;
    v.e3 = i == 3 ? x : v.e3;
}

PT_HD void vol_init(VolumeState& v)
{
    v.e0 = v.e1 = v.e2 = v.e3 = vol_entry(VOL_AIR, 0u, true, true);
    v.pos = 0;
    v.incident = -1;
    v.outgoing = -1;
    v.inside = 0;
    v.distance = 0.0f;
    v.wavelength = 0.0f;
    v.reset = 0;
}

// Pushes the material of a hit. Returns true when the boundary is to be
// skipped: the ray is inside a medium of higher priority.
PT_HD bool vol_push(VolumeState& v, int32_t mat, int32_t priority)
{
    if (v.pos == VOL_STACK_SIZE - 1)
        return false; // full: incident / outgoing / inside keep their values
    const uint32_t m = (uint32_t)mat & VOL_AIR;
    // the medium entered last: topmost, odd parity, of another material
    int last = v.pos;
    for (; last >= 0; last--)
    {
        const uint32_t e = vol_get(v, last);
        if (vol_mat(e) != m && vol_topmost(e) && vol_odd(e))
            break;
    }
    bool odd = true;
    for (int i = v.pos; i >= 0; i--)
    {
        const uint32_t e = vol_get(v, i);
        if (vol_mat(e) == m)
        {
            vol_set(v, i, e & ~VOL_TOPMOST);
            odd = !vol_odd(e);
            break;
        }
    }
    v.inside = odd ? 0 : 1;
    if (v.pos < VOL_STACK_SIZE - 1)
        v.pos++;
    vol_set(v, v.pos, vol_entry(m, (uint32_t)priority, odd, true));
    const uint32_t last_entry = vol_get(v, last);
    if (priority < (int32_t)vol_prio(last_entry))
        return true;
    if (odd)
    {
        v.incident = (int32_t)vol_mat(last_entry);
        v.outgoing = (int32_t)m;
    }
    else
    {
        v.incident = (int32_t)m;
        v.outgoing = (int32_t)vol_mat(last_entry);
    }
    return false;
}

PT_HD void vol_pop(VolumeState& v, bool inside)
{
    const uint32_t top = vol_mat(vol_get(v, v.pos));
    if (v.pos > 0)
        v.pos--;
    if (inside)
    {
        int prev = v.pos;
        for (; prev >= 0; prev--)
            if (vol_mat(vol_get(v, prev)) == top)
                break;
        if (prev >= 0)
            for (int i = prev + 1; i <= v.pos; i++)
                vol_set(v, i - 1, vol_get(v, i));
        if (v.pos > 0)
            v.pos--;
    }
    for (int i = v.pos; i >= 0; i--)
    {
        const uint32_t e = vol_get(v, i);
        if (vol_mat(e) == top)
        {
            vol_set(v, i, e | VOL_TOPMOST);
            break;
        }
    }
}

// A state that a later push can rebuild from the material alone: only the
// air on the stack, nothing travelled, no wavelength. Paths whose state is
// canonical keep none in memory.
PT_HD bool vol_canonical(const VolumeState& v)
{
    return v.pos == 0 && v.distance == 0.0f && v.wavelength == 0.0f;
}

// ---------------------------------------------------------------------------
// Dispersion (src/Device/includes/Dispersion.h:433-535)
// ---------------------------------------------------------------------------
constexpr float MIN_SAMPLE_WAVELENGTH = 360.0f;
constexpr float MAX_SAMPLE_WAVELENGTH = 830.0f;

// Cauchy's equation fitted to the base IOR and the Abbe number (OpenPBR
// translucent base, Dispersion.h:485-501). The reference's constant for
// 587.6^2 is 334777.96 = 578.6^2, so the base IOR is met at 578.6 nm.
PT_HD float dispersion_ior(float abbe_number, float scale, float base_ior, float wavelength)
{
    if (scale == 0.0f)
        return base_ior;
    const float sq_587_6 = 334777.96f;
    const float inv_f_minus_c = 0.00000191038851931481f; // 486.1^-2 - 656.3^-2
    float abbe = abbe_number / scale;
    float B = (base_ior - 1.0f) / (abbe * inv_f_minus_c);
    float A = base_ior - B / sq_587_6;
    return A + B / (wavelength * wavelength);
}

// wavelength_to_RGB_fit (Dispersion.h:433-458): fitted CIE 1931 -> sRGB
// response whose mean over [360, 830] nm is white.
PT_HD f3 wavelength_to_rgb(float w)
{
    f3 c;
    if (w < 463.0f)
    {
        c.x = -1.2776028240727566e-01f / (1.0f + exp_fixed((w - 4.2680623367293401e+02f) / 8.2197460736637176e+00f)) +
              -1.3925673552505122e-11f * exp_fixed((w - 45.0f) / 1.8175459086411596e+01f);
        c.x += 1.2898689750552100e-01f;
    }
    else if (w > 553.0f)
    {
        float t = (w - 6.0625724092824566e+02f) * (1.0f / 2.6577826611702449e+01f);
        c.x = 1.7963649137825513e+01f * (1.0f / 2.6577826611702449e+01f) * exp_fixed(-0.5f * (t * t));
        c.x += 2.5574660155104657e-03f;
    }
    else
        c.x = 0.0f;
    float tg = (w - 5.4209217455705152e+02f) / -2.9598170255834638e+01f;
    c.y = 3.4962267376163049e+02f * exp_fixed(-0.5f * (tg * tg));
    c.y /= w;
    c.z = exp_fixed(3.2987659944421112e+03f + (-2.0975839709372405e+05f / w) - 4.6368268395094020e+02f * log_fixed(w));
    return mk(c.x * 10.0f, c.y * 10.0f, c.z * 10.0f);
}

// sample_wavelength_uniformly (Dispersion.h:469-474) on a draw r in [0, 1).
PT_HD float wavelength_from_draw(float r)
{
    return r * (MAX_SAMPLE_WAVELENGTH - MIN_SAMPLE_WAVELENGTH) + MIN_SAMPLE_WAVELENGTH;
}

} // namespace pt
