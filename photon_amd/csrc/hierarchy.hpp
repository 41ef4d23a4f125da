// hierarchy.hpp -- host-side set-up of the pressure multigrid for one grid.
//
// Connected-aggregation hierarchy with the reference's rules
// (src/sim/pressure_mg.cpp:28-191): level 0 is the structured 7-point grid
// with boundary-masked face coefficients; each coarse unknown is one
// connected piece of a 2x2x2 block of finer unknowns; Galerkin operator with
// piecewise-constant interpolation; coarsen while n > 2000 (and < 12 levels),
// stop when a level no longer shrinks below 0.8x.  For the GPU's dense
// coarsest solve the hierarchy may continue past 2000 unknowns with the same
// aggregation rule until the coarsest holds at most `dense_max` unknowns.
//
// The set-up runs once per geometry change (the reference's mgDirty_), on the
// host; what the GPU needs is laid out here: ELL (column-major) coarse
// operators, child lists for deterministic gather restriction, and the dense
// FP64 inverse of the grounded coarsest operator.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace phg {

struct HostLevel {
    int n = 0, nx = 0, ny = 0, nz = 0;
    std::vector<float> ax, ay, az;          // level 0: masked face coefficients
    std::vector<int> rowStart, col;         // coarse: off-diagonal couplings (positive values)
    std::vector<float> val;
    std::vector<int> bi, bj, bk;            // coarse: block coordinates
    std::vector<int> agg;                   // unknown -> next level unknown (empty on the coarsest)
    std::vector<float> diag;
    std::vector<uint8_t> leak;              // touches an open (Dirichlet) boundary face
    // derived device layouts
    int ellk = 0;                           // max couplings per row
    std::vector<int> ell_col;               // [ellk][n], padded with the row itself
    std::vector<float> ell_val;             // [ellk][n], padded with 0
    std::vector<int> cptr, cidx;            // children of each next-level unknown, ascending
};

struct HostHierarchy {
    std::vector<HostLevel> L;
    int ref_levels = 0;                     // levels produced by the reference's stopping rule
    int nc = 0;                             // coarsest size
    std::vector<double> cinv;               // nc*nc inverse of the grounded coarsest operator (nc <= dense_max)
    std::vector<uint8_t> cmask;             // coarsest rows held at zero (diag <= 0)
    // Level schedule of the reference's Gauss-Seidel sweeps on the coarsest
    // level (pressure_mg.cpp:258-266, coarsest >= level 1): gs_ord holds the
    // rows of the forward sweep grouped by level, then those of the backward
    // sweep; gs_lptr the gs_nf + 1 forward level starts, then the gs_nb + 1
    // backward ones.  A row's level is one past the highest level of the
    // neighbours the sweep updates before it, so the rows of one level are
    // independent and a sweep level by level reads exactly the values the
    // serial sweep reads.
    std::vector<int> gs_ord, gs_lptr;
    int gs_nf = 0, gs_nb = 0;
};

// largest coarsest level the library takes (the Gauss-Seidel sweeps keep
// its vectors in one workgroup's LDS)
constexpr int kCoarseGsMax = 16384;

// wind_closed: 4 flags (x-low, x-high, y-low, y-high) or nullptr.
// The dense coarsest inverse is formed only when the coarsest level has at
// most dense_max unknowns; a coarsest level above kCoarseGsMax is refused.
// Returns false with a message on inconsistent input.
bool build_hierarchy(int nx, int ny, int nz, const float* ax, const float* ay, const float* az,
                     const uint8_t* blocked, const uint8_t* wind_closed, int dense_max, HostHierarchy& H,
                     std::string* err);

// The rest of build_hierarchy once H.L holds levels 0 and 1 (level 0's child
// lists may be there already): the coarser levels, ELL layouts, child lists,
// the coarsest level's Gauss-Seidel schedule and dense inverse.
bool finish_hierarchy(HostHierarchy& H, int dense_max, std::string* err);

// Levels 0 and 1 of build_hierarchy formed on the device, with the same
// results bit for bit: d_ax/d_ay/d_az are the wind-masked level-0
// coefficients and d_blocked the solid mask, in device memory; host copies
// of the masked coefficients (ax/ay/az) go to H.L[0].  Fills H.L[0] (diag,
// leak, agg, child lists) and H.L[1] (operator, diagonal, leak flags, block
// coordinates).  stream: the hipStream_t to work on.  Returns "" or why it
// could not (sizes beyond its key packing); the caller then builds on the host.
std::string device_levels01(void* stream, int nx, int ny, int nz, const float* d_ax, const float* d_ay,
                            const float* d_az, const uint8_t* d_blocked, const float* ax, const float* ay,
                            const float* az, HostHierarchy& H);

// Pressure zones: gas cells joined through faces with a > 0; open when a
// side or top boundary face with a > 0 is reached (src/sim/gasphase.cpp:111-158).
int compute_zones(int nx, int ny, int nz, const float* ax, const float* ay, const float* az, const uint8_t* blocked,
                  std::vector<int>& zone, std::vector<uint8_t>& zone_open);

}  // namespace phg
