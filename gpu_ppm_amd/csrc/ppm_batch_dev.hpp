// Types shared by the batched host driver (ppm_batch.cpp) and its kernels
// (ppm_batch_kernels.hip).
#pragma once

#include <cstddef>
#include <cstdint>

#include "ppm_internal.hpp"

namespace ppm {

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

// A work item: contiguous ring tasks of one spot of one star, and one chunk
// of the energy axis.
struct BatchItem {
  int task0, task1;
  int star, spot;
  int e0;     // first energy of the chunk
  int pad;
};

struct BatchLaunch {
  const SpotParams* spots;
  const RingTask* tasks;
  RingDesc* desc;
  const double* ph;
  const float* fine;
  const double* energies;
  const ClassSet* csets;
  const ClassHdr* chdr;
  const ClassPt* cpts;
  const BatchItem* items;
  const int* order;        // queue order of the items (largest first)
  const int* star_item0;   // [n_stars + 1] item ranges per star
  int* queue;
  unsigned long long* n_evals;  // (position, energy) evaluations performed
  double* acc;             // [n_items][copies][n_phase][n_e]
  double* seg;             // [n_wg][segment totals], workgroup scratch
  double* flux;            // [n_stars][n_phase][n_e]
  int n_items, n_stars, n_e, n_phase, n_fine;
  int nsx_rows, cov_words, ev_cap, atmosphere;
};

size_t batch_lds_bytes(int n_fine, int atmosphere, int nsx_rows, int cov_words, int ev_cap);
int batch_threads();
int batch_acc_copies();  // accumulator copies per item
size_t batch_seg_doubles();  // workgroup scratch of the segment totals
int batch_energies();
int batched_prepare(int atmosphere, size_t lds);  // outside any stream capture
int launch_batched(const DevNsx& N, const BatchLaunch& B, int n_wg, void* stream);
int launch_reduce_batched(const BatchLaunch& B, void* stream);
int batched_occupancy(int atmosphere, size_t lds, int* wg_per_cu);

}  // namespace ppm
