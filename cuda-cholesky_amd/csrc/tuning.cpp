// Block sizes chosen by measurement.
//
// cuda-cholesky_amd/autotune.py times the factorisations over a grid of
// outer block sizes, and the persistent schedule of the batched
// factorisation, on the GPU at hand and writes the winners to a text table,
// one line per entry:
//     potrf_nb <precision> <n_max> <nb>
//     batched_per_cu <precision> <inverse> <workgroups per CU>
//     gemm_tile <family> <min_tiles>
//     panel_launch <precision> <0|1>
//     trailing_cap <precision> <t_max> <workgroups>
// precision: D (FP64), Z (complex FP64), S (FP32 products / storage), B
// (BF16 products).  For an order n the first potrf_nb line of that precision
// with n <= n_max applies.  gemm_tile: a launch of kernel family F takes its
// larger tile when it has at least min_tiles of them (families below).
// panel_launch: 1 factors each panel of the lower real factorisation in one
// launch (panel_potrf), 0 as a chain of diagonal-block and panel-product
// launches.  trailing_cap: a trailing update of the factorisation whose
// order is at most t_max runs on at most that many workgroups (each then
// takes several tiles), so that the look-ahead chain's launches find free
// room on the CUs instead of waiting for an update workgroup to end.  The
// table is read once, from $CUCHOLESKY_TUNING or else tuning.txt beside the
// library; without it the built-in defaults hold.
#include "internal.h"

#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace hc {
namespace {

struct Entry {
  char prec;
  long n_max;
  long nb;
};

std::vector<Entry> &table() {
  static std::vector<Entry> t;
  return t;
}

// batched_per_cu entries, indexed [precision D/S][inverse]; 0 = none
int g_batched_per_cu[2][2] = {{0, 0}, {0, 0}};

// gemm_tile entries by family letter; 0 = none
long g_tile_min[26] = {};

// panel_launch entries by precision letter; -1 = none
int g_panel_launch[26] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,
                          -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};

// trailing_cap entries by precision letter; {0, 0} = none
long g_cap[26][2] = {};

void load() {
  std::string path;
  if (const char *e = std::getenv("CUCHOLESKY_TUNING")) {
    path = e;
  } else {
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&load), &info) != 0 && info.dli_fname != nullptr) {
      path = info.dli_fname;
      const size_t slash = path.rfind('/');
      path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tuning.txt";
    }
  }
  if (path.empty()) return;
  FILE *f = std::fopen(path.c_str(), "r");
  if (f == nullptr) return;
  char key[64], prec[8];
  long n_max = 0, nb = 0;
  char line[256];
  while (std::fgets(line, sizeof line, f) != nullptr) {
    if (line[0] == '#') continue;
    long tmin = 0;
    if (std::sscanf(line, "gemm_tile %7s %ld", prec, &tmin) == 2) {
      if (prec[0] >= 'A' && prec[0] <= 'Z' && tmin > 0) g_tile_min[prec[0] - 'A'] = tmin;
      continue;
    }
    if (std::sscanf(line, "panel_launch %7s %ld", prec, &tmin) == 2) {
      if (prec[0] >= 'A' && prec[0] <= 'Z' && (tmin == 0 || tmin == 1)) g_panel_launch[prec[0] - 'A'] = (int)tmin;
      continue;
    }
    if (std::sscanf(line, "%63s %7s %ld %ld", key, prec, &n_max, &nb) != 4) continue;
    if (std::strcmp(key, "potrf_nb") == 0 && nb > 0 && n_max > 0)
      table().push_back(Entry{prec[0], n_max, nb});
    else if (std::strcmp(key, "trailing_cap") == 0 && prec[0] >= 'A' && prec[0] <= 'Z' && n_max > 0 &&
             nb >= 8) {
      g_cap[prec[0] - 'A'][0] = n_max;
      g_cap[prec[0] - 'A'][1] = nb;
    }
    else if (std::strcmp(key, "batched_per_cu") == 0 && (prec[0] == 'D' || prec[0] == 'S') &&
             (n_max == 0 || n_max == 1) && nb > 0 && nb <= 64)
      g_batched_per_cu[prec[0] == 'S' ? 1 : 0][n_max] = (int)nb;
  }
  std::fclose(f);
}

std::once_flag g_once;

}  // namespace

long tuned_potrf_nb(char prec, long n) {
  std::call_once(g_once, load);
  for (const Entry &e : table())
    if (e.prec == prec && n <= e.n_max) return e.nb;
  return 0;
}

int tuned_batched_per_cu(char prec, bool inverse) {
  std::call_once(g_once, load);
  return g_batched_per_cu[prec == 'S' ? 1 : 0][inverse ? 1 : 0];
}

long tuned_tile_min(char family, long dflt) {
  if (family < 'A' || family > 'Z') return dflt;
  // $CUCHOLESKY_TILE_MIN_<family> (read on every call: the autotuner changes
  // it between measurements), else the table, else the built-in default
  char name[32] = "CUCHOLESKY_TILE_MIN_";
  name[20] = family;
  name[21] = '\0';
  if (const char *e = std::getenv(name)) {
    const long v = std::atol(e);
    if (v > 0) return v;
  }
  std::call_once(g_once, load);
  const long t = g_tile_min[family - 'A'];
  return t > 0 ? t : dflt;
}

bool tuned_panel_launch(char prec, bool dflt) {
  if (prec < 'A' || prec > 'Z') return dflt;
  // $CUCHOLESKY_PANEL_LAUNCH_<precision> (read on every call, for the
  // autotuner), else the table, else the default
  char name[40] = "CUCHOLESKY_PANEL_LAUNCH_";
  name[24] = prec;
  name[25] = '\0';
  if (const char *e = std::getenv(name)) return std::atoi(e) != 0;
  std::call_once(g_once, load);
  const int v = g_panel_launch[prec - 'A'];
  return v < 0 ? dflt : v != 0;
}

int tuned_trailing_cap(char prec, long t) {
  if (prec < 'A' || prec > 'Z') return 0;
  // $CUCHOLESKY_TRAILING_CAP_<precision> = "t_max:workgroups" (read on every
  // call, for the autotuner; "0:0" switches the cap off), else the table
  long t_max = 0, wgs = 0;
  char name[40] = "CUCHOLESKY_TRAILING_CAP_";
  name[24] = prec;
  name[25] = '\0';
  if (const char *e = std::getenv(name)) {
    if (std::sscanf(e, "%ld:%ld", &t_max, &wgs) != 2) t_max = wgs = 0;
  } else {
    std::call_once(g_once, load);
    t_max = g_cap[prec - 'A'][0];
    wgs = g_cap[prec - 'A'][1];
  }
  return t <= t_max && wgs >= 8 ? (int)wgs : 0;
}

}  // namespace hc
