// mpnn.h -- the learned PPE warm start and preconditioner: a three-layer message-passing network
// over the PPE matrix graph, BF16 operands on MFMA with FP32 accumulation.
//
// The graph is the assembled PPE system itself (SELL-64, sorted rows): its active rows are the
// nodes, its off-diagonal entries the edges, with weights alpha_ij = -A_ij / A_ii.  Per layer l
//     u = h U_l,   m_i = sum_j alpha_ij u_j,   z = h W_l + m V_l + c_l,   h' = relu(z)  (l = 1, 2)
// and y = z_3 (two outputs).  Hidden width 128, message width 32, 16 input features.  Outputs:
// x0 = s y[:, 0] (s = RMS of b_i / A_ii over the active rows) and omega = 2 sigmoid(y[:, 1]), the
// weight of the diagonal part of the preconditioner.  Parameters are kept in FP32 (flat, in the
// order U1 W1 V1 c1 U2 W2 V2 c2 U3 W3 V3 c3, row-major) for Adam and packed to BF16 for the MFMA
// kernels.  The reference has no counterpart: it solves no PPE.
#pragma once
#include "mpsr_internal.h"

namespace mpsr {

constexpr int NN_F = 16, NN_H = 128, NN_M = 32, NN_OUT = 2;

// The PPE system as the network sees it (rows in sorted order, FP64 vectors).
struct MpnnSys {
    int N = 0, nslices = 0, rows_active = 0;
    const int* slice_off = nullptr;
    const int* pos = nullptr;      // SELL position of each row
    const int* slice_w = nullptr;
    const int* cols = nullptr;
    const int* cnt = nullptr;      // stored entries per row
    const int* dir = nullptr;      // Dirichlet rows (not nodes of the graph)
    const float* v32 = nullptr;    // off-diagonal values, one of v32 / v64
    const double* v64 = nullptr;
    const double* diag = nullptr;
    const double* rhs = nullptr;
    const double* xprev = nullptr; // previous pressure, sorted (0 on Dirichlet rows)
    const double* xsol = nullptr;  // the plain solver's solution (training target)
};

template <typename R> struct MpnnGeo {
    const R4<R>* spos = nullptr;
    const R4<R>* svel = nullptr;
    const int* perm = nullptr;
    const R* stl_mag = nullptr;    // slot order
    double l0 = 0, dt = 0, re = 0;
};

struct Mpnn;
Mpnn* mpnn_create(int cap_rows);
void mpnn_destroy(Mpnn* n);
int mpnn_param_count();
mpsr_status mpnn_set_weights(Mpnn* n, const float* w, hipStream_t st);
mpsr_status mpnn_get_weights(Mpnn* n, float* w, hipStream_t st);
mpsr_status mpnn_get_grads(Mpnn* n, float* g, hipStream_t st);
bool mpnn_has_weights(const Mpnn* n);
// features of the active rows (X, bf16) and the scale s
template <typename R> void mpnn_features(Mpnn* n, const MpnnSys& sys, const MpnnGeo<R>& geo, hipStream_t st);
void mpnn_forward(Mpnn* n, const MpnnSys& sys, hipStream_t st);
// x0 = s y0 on the active rows (0 elsewhere), omega = 2 sigmoid(y1) clamped to [0.5, 1.5]
void mpnn_apply(Mpnn* n, const MpnnSys& sys, double* x0, double* omega, hipStream_t st);
// loss of the current system against sys.xsol and the gradients of all parameters; parts =
// {total, supervised, residual, preconditioner} (synchronises)
mpsr_status mpnn_loss_grad(Mpnn* n, const MpnnSys& sys, double beta, double* parts, hipStream_t st);
void mpnn_adam(Mpnn* n, float lr, hipStream_t st);
// X [N][16] and Y [N][2] in sorted row order, and s (synchronises)
mpsr_status mpnn_get_rows(Mpnn* n, int N, float* X, float* Y, double* s, hipStream_t st);

}  // namespace mpsr
