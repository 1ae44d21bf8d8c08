// dim_params.hpp — parameter blocks passed by value to the DIM kernels.
#pragma once

namespace dimhip {

enum : int { EOS_IDEAL = 0, EOS_STIFFENED = 1 };
enum : int { LAMBDA_KAPILA = 0, LAMBDA_ALLAIRE = 1, LAMBDA_EQUAL_VELOCITY = 2 };
enum : int { BC_TRANSMISSIVE = 0, BC_REFLECTIVE = 1, BC_NONREFLECTIVE = 2, BC_REFLECTION = 3 };

constexpr int kMaxMat = 4;

// Material constants.  gm1 = gamma - 1 and gpinf = gamma * p_inf are the
// reference's own subexpressions ((params.gamma - 1.0), params.gamma *
// params.p_inf in src/euler/eos.hpp), computed once on the host with the same
// double operations.
struct EosParams {
    int kind[kMaxMat];
    double gamma[kMaxMat];
    double pinf[kMaxMat];
    double gm1[kMaxMat];
    double gpinf[kMaxMat];
    int lambda_model;
    double alpha_source_floor;
};

// Single-precision copies for the mixed-precision form.
struct EosParamsF {
    int kind[kMaxMat];
    float gamma[kMaxMat];
    float pinf[kMaxMat];
    float gm1[kMaxMat];
    float gpinf[kMaxMat];
    float inv_gm1[kMaxMat];
    int lambda_model;
    float alpha_source_floor;
};

// Grid geometry: N[d] cells per axis, x fastest; SoA planes of `pitch`
// elements, field f of cell c at q[f * pitch + c].
struct Grid {
    int N[3];
    long long stride[3];
    long long ncells;
    long long pitch;
    double dx[3];
};

// Per-context step control block, resident in device memory.  `dtmin` holds
// the bit patterns of positive doubles (ordered like unsigned integers) for
// an atomic min; slot `slot` is the current step's CFL minimum, the other
// slot collects the next step's.
struct StepCtl {
    unsigned long long dtmin[2];
    double time;
    double dt;
    double dt_max;
    double t_end;
    long long step;
    int slot;
    int active;
    int use_t_end;
    int err;
};

enum : int { ERR_NONFINITE_HLLC = 1, ERR_NONPOSITIVE_DT = 2 };

}  // namespace dimhip
