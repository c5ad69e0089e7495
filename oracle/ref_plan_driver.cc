// ref_plan_driver.cc - exposes the reference's own planner header (include/plan.hpp of the
// reference tree, compiled from where it lies) through a C ABI so that the oracle's index
// map, cycle count, divisor enumeration and feasibility tests can be checked against it.
// TEST INFRASTRUCTURE ONLY: loaded by the CPU test suite, never by the library.
#define CL_TARGET_OPENCL_VERSION 120
#include <CL/cl.h>
#include <cstring>
#include <vector>
#include "plan.hpp"

extern "C" {

// Transposition::Next (plan.hpp:54-56), 32-bit unsigned arithmetic as in the reference.
unsigned ref_next(unsigned i, int m, int n) {
  Transposition t(m, n);
  return t.Next(i);
}

// Transposition::GetNumCycles (plan.hpp:57-73).
unsigned ref_num_cycles(int m, int n) {
  Transposition t(m, n);
  return t.GetNumCycles();
}

// Factorize::tiling_options (plan.hpp:163-219): tile sizes in the reference's order.
int ref_tile_sizes(int n, int *out, int cap) {
  Factorize f(n);
  f.tiling_options();
  std::vector<int> &v = f.get_tile_sizes();
  int c = (int)v.size();
  for (int i = 0; i < c && i < cap; i++) out[i] = v[i];
  return c;
}

// Feasibility classes (plan.hpp:95-157) with the default GPUInfo (16 KiB local memory,
// 256 work-items; plan.hpp:27-30). kind: 0 = T010_BS(m,n), 1 = T010_PTTWAC(m,n),
// 2 = T0100_PTTWAC(A=m, a=n, B=p, b=q).
int ref_feasible(int kind, int m, int n, int p, int q) {
  switch (kind) {
    case 0: { T010_BS t(m, n); return t.IsFeasible(); }
    case 1: { T010_PTTWAC t(m, n); return t.IsFeasible(); }
    case 2: { T0100_PTTWAC t(m, n, p, q, NULL); return t.IsFeasible(); }
  }
  return -1;
}

}
