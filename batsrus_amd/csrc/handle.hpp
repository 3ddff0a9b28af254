// handle.hpp -- the device-side state behind a bats_handle, shared by the
// C ABI (bats_hip.hip) and the surrogate stage (surrogate.hip).  Internal.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "../../include/bats_hip.h"
#include "stage_kernel.hpp"

using namespace bats;

// buffers of the surrogate stage (allocated on first use)
struct SurrogateState {
  bats_mlp *model;     // network (not owned); null: plain stages
  int mode;            // BATS_SUR_*
  double betaMin;      // gate: minimum plasma beta for the network
  double jumpTol;      // gate: largest scaled jump counted as smooth
  float *face;         // face fluxes, 9 slots per face, all blocks
  __bf16 *rowX;        // network inputs of the gated faces (rows x 16)
  float *rowY;         // network outputs (rows x 8)
  int *rowFace;        // per row: global face index
  float *rowScale;     // per row (float4): density scale, speed scale,
                       // jump norm, face area
  int *counters;       // [0] rows this stage; 64-bit at [2] faces and at
                       // [4] network rows since set_surrogate
  long faceCap;
};

struct bats_handle {
  bats_config cfg;
  int NB;
  long nCellBlk, nCell;
  hipStream_t stream;
  void *U[3];    // state buffers of elemSize-byte reals, vs per variable;
                 // U[2] only for three stages (SSP-RK3)
  long vs;       // variable stride: nCell interior cells + ghost slabs
  int elemSize;  // 8: FP64 state, 4: FP32 state
  int nBuf;      // state buffers: 2, or 3 for three stages
  int cur;       // buffer holding the state at the start of a step
  int stateBuf;  // buffer holding the current State_VGB (differs from cur
                 // between the stages of a 2-stage step)
  double *Usnap;
  void *lowCrit;  // low order face weights (elemSize-byte reals,
                  // 3*NB^2*(NB+1) per block, crit_index order) or null
  int *nei;
  double *cellsize;
  double *dtmaxB;
  double *dtmaxCB;
  double *scal;      // S_NSCAL scalars per run
  int nRun;
  std::vector<int> hRunFirst, hRunOfBlock;
  int *runFirst;     // nRun+1 block offsets
  int *runOfBlock;   // nBlock
  int nWorkgroupMax; // > 0: persistent grid of at most this many workgroups
  int splitShift;    // each block's planes split over 2^splitShift work
                     // items (cells per thread NB >> splitShift)
  int xcdOrder;      // XCD-contiguous order of the work items
  int *err;
  double *partial;
  double *LX, *RX, *LY, *RY, *LZ, *RZ, *FX, *FY, *FZ;
  double *Src;   // Source_VC of the last calc_source block
  int srcBlock;  // that block (0-based), -1: none since the last update
  int iStageBlock;
  std::vector<int> hNei;
  // kernel timing
  std::vector<hipEvent_t> evStart, evStop;
  int nTimed;
  double lastStageMs;
  int lastStageLaunches;
  // bats_advance as one captured graph per time step (global time step and
  // every stage), one per state-buffer parity; captured for tLimit
  int useGraph;
  hipGraphExec_t graphExec[2];
  int graphValid[2];
  double graphTLimit[2];
  // two-level grids: resolution-change slabs (iNeighbor_EB = -2 - slab)
  // and the ghost slabs of fixed outer sides (kind 2) that follow them
  int nSlab, nFine;
  int nResSlab;             // resolution-change slabs (kinds 0 and 1)
  std::vector<int> hSlabs;  // 8 ints per slab
  int *slabs, *slabKind, *fineSlabs;
  void *corr;               // summed fine fluxes of the finer sides
  // the grid as the caller gave it (outer sides -1), kept so that a change
  // of the outer boundary types can rebuild the slab list
  std::vector<int> userNei, userSlabs;
  std::vector<double> userCellSize;
  // outer boundary type of each side (BATS_BC_*) and the ghost state of the
  // fixed/inflow sides, CellState_VI(8, 6) column-major; device copy
  int bcType[6];
  double bcState[6 * NVAR];
  double *dBcState;
  bool anySideRC;           // some block side has DiLevel_EB /= 0
  // B0 splitting (bats_set_b0): face B0 [dir][comp][block][face], cell B0,
  // div B0 and curl B0 [7][cell]; dropped by any grid change
  double *b0F, *b0C;
  int useB0, useB0Src;
  // B0 of the two ghost layers of every outer side (-1 neighbour) [slot]
  // [layer][q][p][3], slot per (block, side) (-1: none), for the fixed,
  // inflow and vary ghost cells
  double *b0Side;
  int *b0SideSlot;
  // time-varying inflow of the x+ side: nVary table points (times and
  // primitive states, 8 per point)
  double *varyT, *varyS;
  int nVary;
  // inner body (bats_set_body): per block the slot of its true cell mask
  // (-1: none), the masks over block + 2-cell halo, the blocks' first cell
  // centres relative to the body centre, the body density and B1rCoef; the
  // work list holds the groups without a body block, then those with one;
  // dropped by any grid change
  int useBody;
  int *bodySlot;
  unsigned char *bodyUsed;
  double *bodyXyz0;
  int *groupList;
  int nPlainGroup, nBodyGroup;
  double bodyRho, b1rCoef;
  // learned face-flux surrogate (surrogate.hip), FP32 state only
  SurrogateState sur;
};

// drop captured step graphs: any change of grid, runs, launch shape, state
// buffers or surrogate invalidates them
inline void drop_graphs(bats_handle *h) {
  for (int p = 0; p < 2; p++) {
    if (h->graphValid[p]) (void)hipGraphExecDestroy(h->graphExec[p]);
    h->graphValid[p] = 0;
  }
}

inline int hip_status(hipError_t e) {
  return e == hipSuccess ? BATS_OK : BATS_EHIP;
}
#define HIPCHK(x)                                                        \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      fprintf(stderr, "bats_hip: %s failed: %s (%s:%d)\n", #x,           \
              hipGetErrorString(e_), __FILE__, __LINE__);                \
      return e_ == hipErrorOutOfMemory ? BATS_ENOMEM : BATS_EHIP;        \
    }                                                                    \
  } while (0)

inline Phys make_phys(const bats_config &c) {
  Phys p;
  p.gamma = c.Gamma;
  p.gm1 = c.Gamma - 1.0;
  p.invgm1 = 1.0 / (c.Gamma - 1.0);
  p.cTiny = c.cTiny;
  p.pMin = c.pMin;
  return p;
}

inline int limiter_code(const bats_config &c) {
  return c.nOrder == 1 ? LIM_FIRST : c.iLimiter;
}

// Whether stage iStage adds |dB|^2/2 to the energy.  UseDbTrickNow is off
// for a one-stage time-accurate run and without the half step scheme
// (src/ModSetParameters.f90:3903-3906); when on, the trick applies at the
// first of two stages and at the single stage of a steady-state run
// (src/ModUpdateState.f90:916-918).
inline int db_trick_at(const bats_config &c, int iStage) {
  const bool timeAccurate = c.IsSteadyState == 0;
  const bool now = c.UseDbTrick && c.UseHalfStep &&
                   !(c.nStage == 1 && timeAccurate);
  return now && ((c.nStage == 2 && iStage == 1) ||
                 (c.nStage == 1 && !timeAccurate));
}

inline int err_from_bits(int bits) {
  if (bits & ERRB_NEGFAST2) return BATS_ENEGFAST2;
  if (bits & ERRB_HLLD) return BATS_EHLLD;
  if (bits & ERRB_NAN) return BATS_ENAN;
  return BATS_OK;
}

// one stage of the update through the surrogate kernels (surrogate.hip)
int sur_stage(bats_handle *h, const StageArgs &a);
// fill the ghost slabs of the state buffer U (two-level grids)
int amr_fill_ghosts(bats_handle *h, const void *U);
