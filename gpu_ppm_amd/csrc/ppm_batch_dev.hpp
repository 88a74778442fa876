// Types shared by the batched host driver (ppm_batch.cpp) and its kernels
// (ppm_batch_kernels.hip).
#pragma once

#include <cstddef>
#include <cstdint>

#include "ppm_internal.hpp"

namespace ppm {

// The batched integral sweeps several phase classes of a ring together, each
// in a slot of its own on a virtual position axis: at most kMaxSlots classes
// whose slots span at most kGroupPositions positions (a class of a ring with
// more pixels is swept alone), with at most ev_cap / (4 kGroupRuns) phase
// points, and only on rings of at most kGroupRuns runs of active pixels
// (others sweep one class at a time).  The groups of a ring follow from its
// pixel count and class table alone (class_groups on the host gives the same
// boundaries), so that a ring split into items by class is cut between
// groups.
constexpr int kGroupPositions = 8192;
constexpr int kMaxSlots = 32;
constexpr int kGroupRuns = 2;

// Phase points of a grid grouped by their fractional offset on a ring of n
// pixels: p n = m + delta.  One ClassSet per pixel count n (index 0: the
// one-element point source, index t: n = 4t).
struct ClassSet {
  int first_class, n_classes;
};
struct ClassHdr {
  double delta;  // shared fractional offset
  int first_pt, n_pts;
};
struct ClassPt {
  int m;     // integer shift
  int bin;   // output phase row
  double w;  // phase-point weight
};

// Per-grid facts of a batched launch: one launch can hold stars on several
// grids (the serving loop packs every queued request into one launch).
struct BatchGrid {
  const double* energies;  // [n_e], ascending
  const ClassSet* csets;   // class tables of the grid's phase points
  const ClassHdr* chdr;
  const ClassPt* cpts;
  const double* ph;        // fine phase tables of the grid's ring tasks, [task - task0][nf1]
  const float* fine;       // [task - task0][3][nf1]
  int n_e, n_phase, n_fine, task0;
};

// A work item: contiguous ring tasks of one spot of one star, and one chunk
// of the energy axis.
struct BatchItem {
  int task0, task1;
  int star, spot;
  int e0;      // first energy of the chunk
  int grid;    // index into BatchLaunch::grids
  int en;      // energies in the chunk (its accumulator row length)
  int c0, c1;  // phase classes [c0, c1) of each ring (a ring split by class is one item per range)
  size_t acc;  // first double of the item's accumulator copies, [copies][n_phase][en]
};
constexpr int kAllClasses = 1 << 30;

struct BatchLaunch {
  const SpotParams* spots;
  const RingTask* tasks;
  RingDesc* desc;
  const BatchGrid* grids;
  const BatchItem* items;
  const int* order;        // queue order of the items (largest first)
  const int* star_item0;   // [n_stars + 1] item ranges per star
  int* queue;
  unsigned long long* n_evals;  // (position, energy) evaluations performed
  double* acc;             // item accumulators
  int n_items;
  int n_stars;             // stars of the launch (all grids)
  int n_fine_max;          // LDS layout: the largest grid of the launch
  int nsx_rows, cov_words, ev_cap, atmosphere;
  int epl;                 // energies per lane: 6, or 2 when every grid has <= 128 energies
  int persistent;          // 1: resident workgroups take items from the atomic queue;
                           // 0: one workgroup per item (workgroup b takes queue position b)
  int emode;               // NsxEnergyMode shared by the launch's stars
  int fine_lds;            // 1: the ring's fine tables are staged in LDS, 0: read from global memory
  unsigned long long* rows_flag;  // bit 1 set when a ring needs more NSX rows than staged
};

size_t batch_lds_bytes(int n_fine, int atmosphere, int nm, int nsx_rows, int cov_words, int ev_cap, bool fine_lds);
int batch_threads();
int batch_acc_copies();  // accumulator copies per item
// Phase-timer totals of the batched integral (a build with PPM_PHASE_TIMERS=1;
// false otherwise): n of the kNumPhases totals and the wavefront count.
bool batch_phase_timers(unsigned long long* cycles, int n, bool reset);
int batch_energies(int epl);  // energies per item
int batched_prepare(int atmosphere, int epl, int emode, bool fine_lds, int lds_max);  // outside any stream capture
int launch_batched(const DevNsx& N, const BatchLaunch& B, int n_wg, void* stream);
int launch_zero_u64(unsigned long long* p, int n, void* stream);  // a kernel, not a memset (graph capture)
// flux[s - s0][b][e] for the stars [s0, s1) of one grid (G, its host copy).
int launch_reduce_batched(const BatchLaunch& B, const BatchGrid& G, int s0, int s1, double* flux, void* stream);
int batched_occupancy(int atmosphere, int epl, int emode, bool fine_lds, size_t lds, int* wg_per_cu);

}  // namespace ppm
