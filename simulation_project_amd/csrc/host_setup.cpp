// host_setup.cpp -- host-only set-up helpers of the C ABI (no GPU calls).
//
//   mpsr_calculate_lambda   CMultiGPUHandler::CalculateLambdaValueLocal
//                           (Source/MultiGPUHandler.cpp:1132-1296)
//   mpsr_auto_compute_zone  CMultiGPUHandler::SetComputeZone(particles, triangles)
//                           (MultiGPUHandler.cpp:111-150)
//   mpsr_make_gridbox       GetSectionComputeZone(zone, (1,1,1), 0,0,0) + CCellSet::SetComputeZone
//                           (solver/Common/inc/Utility_Inline.h:1977-2006, Source/CGPUCellSet.cpp:209-253)
#include <algorithm>
#include <cfloat>
#include <cmath>
#include "../../include/mpsr.h"

extern "C" {

// "if(lo < I_0 <= hi)" in the reference is "(lo < I_0) <= hi": it holds when hi >= 1 or when
// I_0 <= lo, and the first branch for which it holds sets (a, b).
float mpsr_calculate_lambda(float l0) {
    struct Branch { double lo, hi, a, b; };
    static const Branch kFit[] = {
        {0, 0.0001, 5.02608E-06, 3.82946E-10},        {0.0001, 0.0005, 0.000124918, 1.81984E-08},
        {0.0005, 0.001, 0.0017886, -4.84707E-07},      {0.001, 0.002, 0.00373277, -2.15486E-06},
        {0.002, 0.003, 0.0066746, -6.72384E-06},       {0.003, 0.004, 0.0107452, -1.76386E-05},
        {0.004, 0.005, 0.0141993, -3.08052E-05},       {0.005, 0.006, 0.0164975, -4.14937E-05},
        {0.006, 0.007, 0.0181782, -4.95251E-05},       {0.007, 0.008, 0.0205885, -6.28759E-05},
        {0.008, 0.009, 0.0222761, -7.17244E-05},       {0.009, 0.01, 0.0231972, -7.34117E-05},
        {0.01, 0.02, 0.0369119, -0.000210081},         {0.02, 0.03, 0.0667457, -0.000672376},
        {0.03, 0.04, 0.106105, -0.00171939},           {0.04, 0.05, 0.141998, -0.00308074},
        {0.05, 0.06, 0.164981, -0.00414971},           {0.05, 0.06, 0.164981, -0.00414971},
        {0.06, 0.07, 0.178483, -0.00474469},           {0.07, 0.08, 0.197539, -0.00567825},
        {0.08, 0.09, 0.208407, -0.00598107},           {0.09, 0.1, 0.211437, -0.00543137},
        {0.01, 0.5, 0.369118, -0.021008},              {0.5, 1, 1.78861, -0.48471},
        {1, 10, 0.423015, 4.1213},                     {10, 1000, 0.0138964, 745.762},
    };
    float a = 0.0f, b = 0.0f;
    for (const Branch& br : kFit) {
        const double lower_test = ((double)l0 > br.lo) ? 1.0 : 0.0;
        if (lower_test <= br.hi) {
            a = (float)br.a;
            b = (float)br.b;
            break;
        }
    }
    return a * l0 + b;
}

mpsr_status mpsr_auto_compute_zone(const mpsr_params* P, const mpsr_particles* pp, const mpsr_triangle* tr,
                                   int nt, mpsr_box* out) {
    if (!P || !out) return MPSR_PARAMERR;
    mpsr_float3 mn = {FLT_MAX, FLT_MAX, FLT_MAX}, mx = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    auto take = [&](float x, float y, float z) {
        if (x <= mn.x) mn.x = x;
        if (y <= mn.y) mn.y = y;
        if (z <= mn.z) mn.z = z;
        if (x >= mx.x) mx.x = x;
        if (y >= mx.y) mx.y = y;
        if (z >= mx.z) mx.z = z;
    };
    if (pp)
        for (int i = 0; i < pp->n; ++i)
            take((float)pp->position[3 * i], (float)pp->position[3 * i + 1], (float)pp->position[3 * i + 2]);
    if (tr)
        for (int i = 0; i < nt; ++i)
            for (int k = 0; k < 3; ++k) take(tr[i].Vertex[k].x, tr[i].Vertex[k].y, tr[i].Vertex[k].z);
    const float coef = std::max(P->LaplacianInfluenceRadiusCoefficient, P->GradientInfluenceRadiusCoefficient);
    const float off = P->InitialParticleDistance * coef;
    out->m_MaxBound = {mx.x + off, mx.y + off, mx.z + off};
    out->m_MinBound = {mn.x - off, mn.y - off, mn.z - off};
    return MPSR_NOERR;
}

mpsr_status mpsr_make_gridbox(const mpsr_params* P, const mpsr_box* z, mpsr_gridbox* gb) {
    if (!P || !z || !gb) return MPSR_PARAMERR;
    // section (0,0,0) of a (1,1,1) split: Min + Size*0 and Min + Size*1, Size = extent / 1
    const mpsr_float3 size = {(z->m_MaxBound.x - z->m_MinBound.x) / 1.0f, (z->m_MaxBound.y - z->m_MinBound.y) / 1.0f,
                              (z->m_MaxBound.z - z->m_MinBound.z) / 1.0f};
    const mpsr_float3 lo = {z->m_MinBound.x + size.x * 0.0f, z->m_MinBound.y + size.y * 0.0f,
                            z->m_MinBound.z + size.z * 0.0f};
    const mpsr_float3 hi = {z->m_MinBound.x + size.x * 1.0f, z->m_MinBound.y + size.y * 1.0f,
                            z->m_MinBound.z + size.z * 1.0f};
    const float coef = std::max(P->LaplacianInfluenceRadiusCoefficient, P->GradientInfluenceRadiusCoefficient);
    const float r = P->InitialParticleDistance * coef;
    if (!(r > 0.0f)) return MPSR_PARAMERR;
    gb->m_ComputeZone.m_MinBound = lo;
    gb->m_ComputeZone.m_MaxBound = hi;
    gb->m_ParentZone = *z;
    gb->m_PressureZone.m_MaxBound = {hi.x + r, hi.y + r, hi.z + r};
    gb->m_PressureZone.m_MinBound = {lo.x - r, lo.y - r, lo.z - r};
    gb->m_BufferedZone.m_MaxBound = {hi.x + 2 * r, hi.y + 2 * r, hi.z + 2 * r};
    gb->m_BufferedZone.m_MinBound = {lo.x - 2 * r, lo.y - 2 * r, lo.z - 2 * r};
    gb->m_CellSize = r;
    const float gx = (hi.x - lo.x) / r + 4, gy = (hi.y - lo.y) / r + 4, gz = (hi.z - lo.z) / r + 4;
    gb->m_GridSize = {(int)std::ceil(gx), (int)std::ceil(gy), (int)std::ceil(gz)};
    if (!(gb->m_GridSize.x > 0 && gb->m_GridSize.y > 0 && gb->m_GridSize.z > 0)) return MPSR_PARAMERR;
    const double cells = (double)gb->m_GridSize.x * gb->m_GridSize.y * gb->m_GridSize.z;
    if (cells > 2.0e9) return MPSR_PARAMERR;
    return MPSR_NOERR;
}

}  // extern "C"
