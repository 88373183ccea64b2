// fr_cnn.h -- the learned orientation proposal: network shape, parameter
// layout and launchers (fr_cnn.hip).
#pragma once
#include "fr_common.h"
#include "fr_kernels.h"

namespace fr {

// A stack of `nconv` 3x3 / stride 2 / pad 1 convolutions with ReLU
// (1 -> C -> C ... channels, S -> S/2 -> ... pixels) and one fully connected
// layer to K logits.  Parameters are one flat FP32 array: for each
// convolution W[C_out][C_in][3][3] then b[C_out], then W_fc[K][F], b_fc[K]
// with F = C * (S >> nconv)^2.
struct CnnLayout {
  int S = 0, C = 0, L = 0, K = 0;
  long long w_off[8] = {}, b_off[8] = {};
  long long fw_off = 0, fb_off = 0, total = 0;
  int cin(int l) const { return l == 0 ? 1 : C; }
  int hin(int l) const { return S >> l; }   // input side of convolution l
  int hout(int l) const { return S >> (l + 1); }
  int F() const { return C * (S >> L) * (S >> L); }
};

// 0 on success; fills the layout for a valid shape.
int cnn_layout(int S, int C, int L, int K, CnnLayout* out);

// workspace (floats) that forward (+ backward) needs for a batch of B
size_t cnn_workspace(const CnnLayout& lo, int B, bool backward);

int cnn_init(Ctx& c, const CnnLayout& lo, unsigned long long seed, float* master, float* wq);
int cnn_forward(Ctx& c, const CnnLayout& lo, const float* wq, int B, const float* img, float* ws,
                float* logits);
// inference with BF16 activations (C = 64): workspace in bytes
bool cnn_bf16_ok(const CnnLayout& lo);
size_t cnn_workspace_bf16(const CnnLayout& lo, int B);
int cnn_forward_bf16(Ctx& c, const CnnLayout& lo, const float* wq, int B, const float* img,
                     void* ws, float* logits);
// bf16: every GEMM of the step (forward included) with BF16 operands and
// FP32 accumulation (the mixed-precision training step); else FP32 throughout
int cnn_backward(Ctx& c, const CnnLayout& lo, const float* wq, int B, const float* img,
                 const int* label, float* ws, float* loss, float* grad, bool bf16 = false);
int cnn_adam(Ctx& c, const CnnLayout& lo, float* master, float* wq, const float* grad, float* m,
             float* v, int step, float lr);
int cnn_propose(Ctx& c, int K, int B, const float* logits, int topk, int* top, float* prob);
int cnn_images(Ctx& c, int nsam, int S, int B, const float* datc, const float* ctff, float* buf,
               float* img);
int cnn_synth(Ctx& c, int nsam, int S, int B, const float* a3df, const float* dm,
              const float* spar, const PrepGlobals& g, float disc_pixels,
              unsigned long long seed, float* buf, float* img);
int orient_nearest(Ctx& c, int n, const float* dm, int ncls, const float* dmc, int nsym,
                   const float* sym, int* label);
int orient_within(Ctx& c, int na, const float* dma, int nb, const float* dmb, int nsym,
                  const float* sym, float cos_min, unsigned* mask);

}  // namespace fr
