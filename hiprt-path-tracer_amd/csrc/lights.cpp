// lights.cpp -- host side of emissive-triangle sampling by power: the Vose
// alias table (Utils::compute_alias_table, src/Utils/Utils.cpp:63-131) over
// the emitted power of each emissive triangle (GPURenderer.cpp:221-295).
// Compiled with -ffp-contract=off.
#include <cmath>
#include <deque>
#include <vector>

#include "internal.h"

namespace mipt {

void compute_alias_table(const float* input, int32_t n, float in_total_sum, float* out_probas, int32_t* out_alias)
{
    if (n <= 0)
        return;
    // Elements normalised in FP64 so that their mean is 1.
    const double total = in_total_sum;
    std::vector<double> scaled((size_t)n);
    for (int32_t i = 0; i < n; i++)
    {
        scaled[i] = (double)input[i] / total;
        scaled[i] *= (double)n;
    }
    for (int32_t i = 0; i < n; i++)
    {
        out_probas[i] = 0.0f;
        out_alias[i] = 0;
    }
    std::deque<int32_t> small_q, large_q;
    for (int32_t i = 0; i < n; i++)
    {
        if (scaled[i] < 1.0)
            small_q.push_back(i);
        else
            large_q.push_back(i);
    }
    while (!small_q.empty() && !large_q.empty())
    {
        const int32_t s = small_q.front();
        const int32_t l = large_q.front();
        small_q.pop_front();
        large_q.pop_front();
        out_probas[s] = (float)scaled[s];
        out_alias[s] = l;
        scaled[l] = (scaled[l] + scaled[s]) - 1.0;
        if (scaled[l] > 1.0)
            large_q.push_back(l);
        else
            small_q.push_back(l);
    }
    // Whatever is left holds (up to rounding) exactly its own share.
    for (int32_t i : large_q)
        out_probas[i] = 1.0f;
    for (int32_t i : small_q)
        out_probas[i] = 1.0f;
}

void build_power_alias_table(Ctx& ctx)
{
    RenderState& R = ctx.render;
    const int32_t E = ctx.emissive_count;
    R.alias_probas_host.assign((size_t)E, 0.0f);
    R.alias_alias_host.assign((size_t)E, 0);
    R.alias_sum = 0.0f;
    if (E == 0)
    {
        R.alias_probas.release();
        R.alias_alias.release();
        return;
    }
    std::vector<int32_t> emissive((size_t)E);
    HIP_CHECK(hipMemcpy(emissive.data(), ctx.emissive_prims.ptr, sizeof(int32_t) * (size_t)E, hipMemcpyDeviceToHost));
    std::vector<int32_t> mat_idx((size_t)ctx.triangle_count);
    HIP_CHECK(hipMemcpy(mat_idx.data(), ctx.material_indices.ptr, sizeof(int32_t) * (size_t)ctx.triangle_count,
                        hipMemcpyDeviceToHost));
    std::vector<float> power((size_t)E);
    float sum = 0.0f;
    for (int32_t i = 0; i < E; i++)
    {
        const int32_t t = emissive[i];
        const float* v[3];
        for (int k = 0; k < 3; k++)
            v[k] = &ctx.host_vertices[3 * (size_t)ctx.host_indices[3 * (size_t)t + k]];
        const float ab[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
        const float ac[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
        const float nx = ab[1] * ac[2] - ab[2] * ac[1];
        const float ny = ab[2] * ac[0] - ab[0] * ac[2];
        const float nz = ab[0] * ac[1] - ab[1] * ac[0];
        const float area = 0.5f * sqrtf(nx * nx + ny * ny + nz * nz);
        const mipt_material& m = ctx.materials_host[mat_idx[t]];
        const float lum = 0.3086f * m.emission[0] + 0.6094f * m.emission[1] + 0.0820f * m.emission[2];
        const float emission_luminance = lum * m.emission_strength * m.global_emissive_factor;
        power[i] = emission_luminance * area;
        sum += power[i];
    }
    compute_alias_table(power.data(), E, sum, R.alias_probas_host.data(), R.alias_alias_host.data());
    R.alias_sum = sum;
    R.alias_probas.reserve(sizeof(float) * (size_t)E);
    R.alias_alias.reserve(sizeof(int32_t) * (size_t)E);
    HIP_CHECK(hipMemcpy(R.alias_probas.ptr, R.alias_probas_host.data(), sizeof(float) * (size_t)E,
                        hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(R.alias_alias.ptr, R.alias_alias_host.data(), sizeof(int32_t) * (size_t)E,
                        hipMemcpyHostToDevice));
}

} // namespace mipt
