// dim_params.hpp — parameter blocks passed by value to the DIM kernels.
#pragma once

namespace dimhip {

enum : int { EOS_IDEAL = 0, EOS_STIFFENED = 1, EOS_NOBLE_ABEL = 2 };
enum : int { LAMBDA_KAPILA = 0, LAMBDA_ALLAIRE = 1, LAMBDA_EQUAL_VELOCITY = 2 };
enum : int { BC_TRANSMISSIVE = 0, BC_REFLECTIVE = 1, BC_NONREFLECTIVE = 2, BC_REFLECTION = 3 };

constexpr int kMaxMat = 4;

// Material constants.  gm1 = gamma - 1 and gpinf = gamma * p_inf are the
// reference's own subexpressions ((params.gamma - 1.0), params.gamma *
// params.p_inf in src/euler/eos.hpp), computed once on the host with the same
// double operations.  covol is the noble-Abel co-volume b (params.covolume).
struct EosParams {
    int kind[kMaxMat];
    double gamma[kMaxMat];
    double pinf[kMaxMat];
    double gm1[kMaxMat];
    double gpinf[kMaxMat];
    double covol[kMaxMat];
    // an upper bound of the minimum internal energy density per unit volume
    // fraction, e_k(p = 1e-10) rho_k: (1e-10 + gamma p_inf) / (gamma - 1) for a
    // stiffened gas and 1e-10 / (gamma - 1) otherwise (noble-Abel: co-volume
    // >= 0), raised by a relative 1e-13 over the rounding of the host's
    // division; +inf where no such bound holds.  repair_state uses it to skip
    // an energy floor that cannot act.
    double emin_ub[kMaxMat];
    int noble_abel;  // nonzero when any material is noble-Abel
    int lambda_model;
    double alpha_source_floor;
    // EosParams compiles the ideal and stiffened-gas laws only; EosParamsNA
    // (same layout) adds the noble-Abel branches, chosen per material by kind
    static constexpr bool kNA = false;
};

struct EosParamsNA : EosParams {
    static constexpr bool kNA = true;
    EosParamsNA() = default;
    explicit EosParamsNA(const EosParams& p) : EosParams(p) {}
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

// Element `off` (a 32-bit count) of a plane whose pointer is uniform across
// the wave: the byte offset is formed in 32 bits, so the access can take the
// scalar base + 32-bit vector offset addressing mode (no 64-bit address per
// lane).
template<typename T>
__device__ __forceinline__ const T& plane_at(const T* plane, unsigned off)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(plane) + off * (unsigned)sizeof(T));
}
template<typename T>
__device__ __forceinline__ T& plane_at(T* plane, unsigned off)
{
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(plane) + off * (unsigned)sizeof(T));
}

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
    // the dt-slot parity a step's fused CFL epilogue writes (slot ^ 1 when
    // the step began; the sweeps read this, not slot, which a fused last
    // sweep flips while it runs)
    int cfl_parity;
    int active;
    int use_t_end;
    int err;
    // the first failing HLLC face in the reference's serial order (the one
    // whose require_positive would throw first): the smallest hllc_err_key
    unsigned long long err_key;
};

// Device error bits.  The HLLC ones follow the reference's require_positive
// checks in their order (hllc.hpp:88-91): a face reports the first of them
// its inputs fail, and the host names the first bit set in that order.
enum : int {
    ERR_HLLC_LEFT_DENSITY = 1,
    ERR_NONPOSITIVE_DT = 2,
    ERR_HLLC_RIGHT_DENSITY = 4,
    ERR_HLLC_LEFT_PRESSURE = 8,
    ERR_HLLC_RIGHT_PRESSURE = 16
};

// Order key of a failing HLLC face: sweep direction, then the reference's
// line index (sweep_core.hpp:173-185), then the face along the line
// (sweep_core.hpp:84-92), then which input check failed (hllc.hpp:88-91).
__device__ __forceinline__ unsigned long long hllc_err_key(int dir, long long line, int face, int err_bit)
{
    const unsigned check = err_bit == ERR_HLLC_LEFT_DENSITY ? 0u : err_bit == ERR_HLLC_RIGHT_DENSITY ? 1u
                           : err_bit == ERR_HLLC_LEFT_PRESSURE ? 2u : 3u;
    return ((unsigned long long)dir << 59) | ((unsigned long long)line << 22) | ((unsigned long long)face << 2) | check;
}

// The CFL minimum of a launch is collected over kDtSlots addresses one
// 128-byte line apart (block b writes slot b % kDtSlots), so that thousands
// of blocks' atomics do not serialise on one address; the step's first
// kernel reduces the slots into StepCtl::dtmin.  The slots of both parities
// follow the StepCtl in the same allocation of kCtlAllocBytes.
constexpr int kDtSlots = 64;
constexpr int kDtStride = 16;  // unsigned long longs per 128-byte line
constexpr size_t kCtlBytes = 256;
constexpr size_t kCtlAllocBytes = kCtlBytes + 2 * kDtSlots * kDtStride * sizeof(unsigned long long);
static_assert(sizeof(StepCtl) <= kCtlBytes, "StepCtl outgrew its header");

__host__ __device__ inline unsigned long long* dt_slots(StepCtl* ctl, int parity)
{
    return (unsigned long long*)((char*)ctl + kCtlBytes) + (size_t)parity * kDtSlots * kDtStride;
}

}  // namespace dimhip
