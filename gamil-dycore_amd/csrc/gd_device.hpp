// gd_device.hpp -- data layout shared by the host orchestration and the HIP
// kernels of the barotropic step.
//
// HBM layout (per member, one arena): every 2-D field is `M` rows of `pitch`
// elements, longitude contiguous (pitch = N rounded up to 32 elements so each
// row starts on a 256-byte boundary).  Half-lat fields use rows 0..M-2.
// State slots:
//   X  - state at the start of a step: gd, U, V and u, v ("full" state)
//   Y,Z- outputs of integrator calls (ping-pong): gd, U, V; u and v are
//        derived on the fly exactly as update_state computes them
//        (dycore_mod.F90:636-645)
//   P1,P2 - the two predictor states of predict_correct
// Tendencies Tb (second predictor) and Tc (corrector), FP64 in every
// precision mode, feed the beta inner products.  In the MFMA filter modes the
// projected momentum rows are staged at off_sy beside the rows at off_sx.
#pragma once

#include <cstdint>

namespace gd {

constexpr int kBlock = 256;   // threads of the row kernels
constexpr int kXcd = 8;       // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr int kSeg = kBlock - 2;  // longitudes per row tile: with the two halo points
                                  // every thread loads exactly one LDS slot

enum Slot { SX = 0, SY = 1, SZ = 2, SP1 = 3, SP2 = 4, NSLOT = 5 };
enum Fld { FGD = 0, FU = 1, FV = 2, Fu = 3, Fv = 4, NFLD = 5 };
enum Comp { CDU = 0, CDV = 1, CDGD = 2 };
enum Phase { PH_A = 0, PH_B = 1, PH_C = 2, PH_T = 3, PH_D = 4, PH_P = 5 };

constexpr int kMaxRad = 24;      // radix passes of the filter FFT plan
constexpr int kMaxFarRng = 4;    // row ranges of a fused phase launch's row tiles

// Entries k = 0..N/2 of the filter's twiddle table (the others are conjugates).
__host__ __device__ inline int tw_entries(int N) { return N / 2 + 1; }
// The table as stored (device and LDS): padded with zeros to whole waves of
// 64 entries, so a wave copies it straight to LDS with no lane out of range.
__host__ __device__ inline int tw_padded(int N) { return (tw_entries(N) + 63) / 64 * 64; }

// Per full row j (0-based).  Half-row entries at index j describe half row j
// (0 for j = M-1).
struct RowCoef {
  double fcos, f, c, fdlon, fdlat;     // full row j (pole rows: reset cos)
  double hcos, hdlon, hdlat;           // half row j
  double k_uadv_lon, k_uadv_lat;       // 0.25/fdlon, 0.25/fdlat
  double k_vadv_lon, k_vadv_lat;       // 0.25/hdlon, 0.25/hdlat
  double c1, c2;                       // hcos[j-1]/fcos[j], hcos[j]/fcos[j]
  double fdlon2, fdlat2, hdlon2, hdlat2;
  double mass_w;                       // (fcos*dlon)*dlat
  int32_t stage[3];                    // staging slot of du/dv/dgd, -1 if none
  int32_t cut[3];                      // their cutoff wavenumbers
};

// Per-member arena.  Members sit `mstride` bytes apart in one allocation and
// every buffer is at a fixed offset from the member's base, so kernels form
// pointers arithmetically from kernel arguments (no table in HBM to load):
//   k * fs, k = slot * NFLD + field   state fields (double)
//   25 fs ghs, 26 fs Tb dgd, 27 fs Tc dgd, 28 fs div, 29 fs vor (double)
//   30 fs + q cs                      Tb du, Tb dv, Tc du, Tc dv (double)
//   off_part, off_dpart, off_sx, off_beta  beta partials [M][3][nseg][2], diag
//                                     partials [M][nseg][4], staged filter rows
//                                     [nslots][pitch], beta of the last call
struct Member {
  char *b;
  size_t fs, cs, off_part, off_dpart, off_sx, off_beta, off_sy;
  double *diag_out;  // [cap][3]: mass, energy, beta per step
  __host__ __device__ double *st(int slot, int f) const { return (double *)(b + (size_t)(slot * NFLD + f) * fs); }
  __host__ __device__ double *ghs() const { return (double *)(b + (size_t)(NSLOT * NFLD) * fs); }
  __host__ __device__ double *tb(int c) const {
    return (double *)(c == CDGD ? b + (size_t)(NSLOT * NFLD + 1) * fs : b + (size_t)(NSLOT * NFLD + 5) * fs + (size_t)c * cs);
  }
  __host__ __device__ double *tc(int c) const {
    return (double *)(c == CDGD ? b + (size_t)(NSLOT * NFLD + 2) * fs : b + (size_t)(NSLOT * NFLD + 5) * fs + (size_t)(2 + c) * cs);
  }
  __host__ __device__ double *div() const { return (double *)(b + (size_t)(NSLOT * NFLD + 3) * fs); }
  __host__ __device__ double *vor() const { return (double *)(b + (size_t)(NSLOT * NFLD + 4) * fs); }
  __host__ __device__ double *part() const { return (double *)(b + off_part); }
  __host__ __device__ double *dpart() const { return (double *)(b + off_dpart); }
  __host__ __device__ double *sx() const { return (double *)(b + off_sx); }
  __host__ __device__ double *beta() const { return (double *)(b + off_beta); }
  // projected momentum rows of the MFMA filter [nslots][pitch] (MFMA modes)
  __host__ __device__ double *sy() const { return (double *)(b + off_sy); }
};

struct StageSlot {
  int32_t comp;   // Comp
  int32_t row;    // full row (du, dgd) or half row (dv)
  int32_t kept;   // kept FFTPACK slots (2c+2)
  int32_t pad;
};

// One staged row of one member, filtered by one workgroup, with everything
// the filter kernel needs to know about it.
struct FiltRow {
  int32_t member;
  int32_t slot;     // staging slot
  int32_t comp;     // CDU / CDV / CDGD
  int32_t row;      // full row (du, dgd) or half row (dv)
  int32_t cut;      // cutoff wavenumber c
  int32_t pad;
  double wrow;      // area weight of the row in the beta inner products
};

struct Args {
  char *arena;                 // member m at arena + m * mstride
  size_t mstride, fs, cs, off_part, off_dpart, off_sx, off_beta, off_sy;
  double *diag;                // [nmem][diag_cap][3]
  const RowCoef *coef;
  const StageSlot *slots;
  const FiltRow *frows;
  const int32_t *mslots;       // k_filter_mm: the staging slots of the momentum rows
  const void *mm_phi;          // k_filter_mm: the basis matrices (mm_basis_fill)
  int mm;                      // k_filter: momentum rows come projected in sy (MFMA modes)
  const double *twid;          // [N/2+1][2]: cos, sin of 2 pi k / N
  const double *ghost;         // [2][pitch]: a row of 1.0 and a row of 0.0, read for rows beyond the grid
  double *sx_ext;              // k_filter projection (PH_P): staged rows at sx_ext + slot * pitch
                               // instead of the member's staging rows (nullptr: the member's)
  int fft_L;                   // complex FFT length: N/2 (even N) or N
  int fft_half;                // 1: even N, real FFT through a half-length complex FFT
  int fft_nrad;                // radix passes
  int fft_rad[kMaxRad];
  int32_t *step_ctr;
  int N, M, pitch, nmem;
  int m0;                      // first member of a row launch
  int nseg;                    // longitude segments of kSeg points per row
  int ntiles;                  // work items of this launch
  // k_phase: nnear whole-row filter tiles (one per FiltRow of the active
  // members; blocks 0..nnear_pad-1), then row tiles over the rows with a
  // component outside the filter, given as nfar_rng ranges of far_cnt[k]
  // rows from far_lo[k] (nfar_rows rows in all)
  int nnear, nnear_pad;
  int nfar_rng, nfar_rows;
  int far_lo[kMaxFarRng], far_cnt[kMaxFarRng];
  int in_full;                 // k_diffuse: read stored u, v of s_in
  int nslots;
  int upd_wpm, nmem_launch;    // k_update: workgroups per member, members
  int upd_narrow;              // k_update: four points and partials per thread per round (else eight)
  int lap_src, lap_final;      // k_diffuse: operand (0 state s_in, 1 slot s_lap), last application
  int s_lap, s_lapout;         // k_diffuse: operand and result slots of intermediate applications
  int isp_mode, upd_isp;       // k_isp mode; k_update applies the isp corrector
  double isp_scale;            // k_isp: 2/subcycles
  int pass, phase;
  int s_in, s0, s_out;
  int upwind;
  int qcon;
  int diag_cap;
  double hdt;                  // dt/2 of a predictor phase
  double dt;                   // integrator time step (corrector)
  double kdiff;                // (sign*dt)*diffusion_coef
  double beta_lon, beta_lat;
  double dlon, dlat;
};

__host__ __device__ inline Member member_at(const Args &a, int m) {
  Member mb;
  mb.b = a.arena + (size_t)m * a.mstride;
  mb.fs = a.fs;
  mb.cs = a.cs;
  mb.off_part = a.off_part;
  mb.off_dpart = a.off_dpart;
  mb.off_sx = a.off_sx;
  mb.off_beta = a.off_beta;
  mb.off_sy = a.off_sy;
  mb.diag_out = a.diag + (size_t)m * a.diag_cap * 3;
  return mb;
}

}  // namespace gd
