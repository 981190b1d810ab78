// Hilbert-curve keys of DofsManager::reorderWithHilbert, shared by the host
// library (pre() on the CPU) and the device set-up (pre() on the GPU).
//
// Reference: fem-master/src/reorder/Hilbert.cpp:17-128 (gray-code entry and
// direction sequences), :225-341 (key of a point), :228-271 (float
// quantisation of the coordinates); dofs/DofsManager.cpp:118-239 (caller).
//
// The curve is held as a state machine: state = direction << dim | entry
// point, one digit (dim coordinate bits) per level from the top bit down.
#pragma once

#include <cstdint>

namespace gfem {

struct HilbertCurve {
  int dim = 0;
  uint8_t nextState[32][8] = {};
  uint8_t digitOut[32][8] = {};

  explicit HilbertCurve(int d) : dim(d)
  {
    const int N = 1 << d;
    const unsigned mask = (1u << d) - 1u;
    auto gray = [](unsigned i) { return i ^ (i >> 1); };
    auto grayInv = [d](unsigned g) {
      unsigned v = g;
      for(int j = 1; j < d; ++j) v ^= g >> j;
      return v;
    };
    auto rotr = [d, mask](unsigned k, int r) { return ((k >> r) | (k << (d - r))) & mask; };
    auto rotl = [d, mask](unsigned k, int r) { return ((k << r) | (k >> (d - r))) & mask; };
    unsigned entry[8], dir[8];
    entry[0] = 0;
    for(int i = 1; i < N; ++i) entry[i] = gray(2u * ((i - 1) / 2));
    dir[0] = 0;
    for(int i = 1; i < N / 2; ++i) dir[i] = ((i % 2 == 0) ? gray(i - 1) : gray(i)) % d;
    for(int i = N / 2; i < N; ++i) dir[i] = dir[N - i - 1];
    for(int dd = 0; dd < d; ++dd)
      for(int e = 0; e < N; ++e) {
        const int s = (dd << d) | e;
        for(int l = 0; l < N; ++l) {
          const unsigned w = grayInv(rotr((unsigned)l ^ (unsigned)e, dd + 1));
          digitOut[s][l] = (uint8_t)w;
          const unsigned ne = (unsigned)e ^ rotl(entry[w], dd + 1);
          const unsigned nd = (dd + dir[w] + 1) % d;
          nextState[s][l] = (uint8_t)((nd << d) | ne);
        }
      }
  }

  uint64_t index(const uint32_t *u, int top) const
  {
    uint64_t h = 0;
    unsigned s = 0;
    for(int m = top; m >= 0; --m) {
      unsigned l = 0;
      for(int k = 0; k < dim; ++k) l |= ((u[k] >> m) & 1u) << k;
      h = (h << dim) ^ digitOut[s][l];
      s = nextState[s][l];
    }
    return h;
  }
};

// Coordinate quantisation (Hilbert.cpp:228-271): bits = 64/dim per axis,
// u = (uint32)(f*scale + shift) with scale = 2^(bits-1)/(max-min) and
// shift = -min*scale, in float, product and sum rounded separately.
struct HilbertQuant {
  int bits = 0;
  float scale[3] = {0.f, 0.f, 0.f}, shift[3] = {0.f, 0.f, 0.f};
  HilbertQuant(int dim, const float *mn, const float *mx)
  {
    bits = 64 / dim;
    const unsigned two = 1u << (bits - 1);
    for(int i = 0; i < dim; ++i) {
      scale[i] = (float)two / (mx[i] - mn[i]);
      shift[i] = -mn[i] * scale[i];
    }
  }
  uint32_t operator()(int i, float f) const
  {
    volatile float t = f * scale[i];
    return (uint32_t)(t + shift[i]);
  }
};

}  // namespace gfem
