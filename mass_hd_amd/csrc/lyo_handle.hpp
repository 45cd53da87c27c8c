// lyo_handle.hpp -- the run handle behind mass_lyo_t (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/mass_lyo.h"
#include "lyo_engine.hpp"

constexpr int kMaxPart = 4096;

struct mass_lyo_s {
  mass_lyo_params prm{};
  int N = 0, NP = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  std::unique_ptr<lyo::Engine> eng;
  double* d_stage = nullptr;  // N*9 doubles (AoS populations) + 3N
  double* d_part = nullptr;   // 2*kMaxPart partial sums + 2 results
  lyo::LaunchCfg launch;
  lyo::DiagScratch diag;
  std::vector<void*> diag_allocs;
  bool has_state = false;     // configured or set_state called
  std::mt19937 gen;           // the run's generator (src/random.cpp:12), seeded by configure
  bool growth = false;        // a go-or-grow run (mass_lyo_set_growth)
  // surrogate corrector
  struct Surrogate {
    mass_lyo_net_t net = nullptr;
    int judge_every = 50;
    double tol = 1e-3;
    long long count = 0;      // steps taken in surrogate mode
    int plain_left = 0;       // plain steps still owed after a failed judgement
    int fused = 1;            // network part in one kernel (FP32, LX and LY >= 16)
    float* delta = nullptr;   // [N][8]
    double* part = nullptr;   // judge partial sums
    // output calibration (mass_lyo_surrogate_set_calibration): decayed sums
    // of t.p and p.p per output over the judged steps, the factors, and the
    // calibrated output scales (host copy and device)
    bool calib = false;
    double ctp[5] = {}, cpp[5] = {};
    float alpha[5] = {1.f, 1.f, 1.f, 1.f, 1.f};
    float h_oscale[8] = {};
    float* d_oscale = nullptr;
    mass_lyo_surrogate_stats stats{};
  } sur;
  std::string err;
};
