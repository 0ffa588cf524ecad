// Poke windows of a poke set: shared between the deformable-mirror kernels (dm.hip), which
// evaluate the poked field changes, and the C ABI (capi.hip), which runs the transforms.
#pragma once

#include "common.hpp"
#include "../../include/hcipy_amd.h"

namespace hcipy {

// Host-side description of the poke windows of one poke-set item.
struct PokeWinItem {
  hcipy_dm_gaussian_desc dm;  // influence functions on the item's pupil grid
  const void* field;          // base pupil field (complex64; complex128 for FP64)
  double phase_scale;         // phase of the poke: 2 k h
  int prec;                   // Prec of mft_gemm.hpp (operand format of the window rows)
  int wx, wy;                 // window width and height in pixels (multiples of 64, at most kPokeWindowMax)
  int64_t row0;               // first global window row of this item
  const int* org;             // device array: window origin (x0, y0) per actuator
  void* out[3];               // [actuator][wy][2 wx] values of each operand array
  float* row_scale;           // [actuator][wy] (split formats)
};

constexpr int kPokeWindowMax = 1024;

int check_gaussian_desc(const hcipy_dm_gaussian_desc* d);
void gaussian_support_window(const hcipy_dm_gaussian_desc& d, int64_t a, int64_t* i0, int64_t* i1, int64_t* j0,
                             int64_t* j1);
int poke_window_bytes(int n_items);
int upload_poke_items(const PokeWinItem* items, int n, void* host, void* dev_items, hipStream_t s);
int launch_poke_windows(const void* dev_items, int n_items, int64_t total_rows, int max_wx, hipStream_t s);

}  // namespace hcipy
