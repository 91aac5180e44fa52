// Learned SPAI preconditioner: kNN geometry features, MLP (FP32) forward/backward/Adam, SPAI application.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cdj {

static constexpr int GEO = 16;      // per-triangle geometry record (floats)
static constexpr int NF = 10;       // features per neighbour
static constexpr int KNN_MAX = 32;  // largest supported neighbourhood

struct Mlp {
  int k = 16, F = 160, H = 128, O = 144;
  int t = 0;  // Adam step
  float* params = nullptr;
  float* grad = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  int64_t nparams() const;
  void offsets(int64_t off[6]) const;
};

struct MeshFeat {
  int64_t n = 0;
  int k = 0;
  float* geo = nullptr;     // n x GEO
  int32_t* nbr = nullptr;   // n x k
  float* feat = nullptr;    // n x (k NF)
};

hipError_t mlp_init(Mlp& m, int k, int hidden, uint64_t seed);
// BF16 inference: weights and activations rounded to BF16, products accumulated in FP32 on MFMA; wbuf holds
// mlp_bf16_weight_count(m) BF16 values.
int64_t mlp_bf16_weight_count(const Mlp& m);
hipError_t mlp_forward_bf16(const Mlp& m, const float* feat, int64_t n, float* acts, unsigned short* wbuf,
                            hipStream_t st);
void mlp_free(Mlp& m);
hipError_t mesh_prepare(const double* P1, const double* P2, const double* P3, int64_t n, int k, MeshFeat& f,
                        hipStream_t st);
void mesh_free(MeshFeat& f);
// acts: n x (2H + O) floats: h1 | h2 | Y
hipError_t mlp_forward(const Mlp& m, const float* feat, int64_t n, float* acts, hipStream_t st);
hipError_t mlp_backward(const Mlp& m, const float* feat, int64_t n, const float* acts, const float* dY, float* work,
                        hipStream_t st);
hipError_t mlp_adam(Mlp& m, float lr, hipStream_t st);
hipError_t jacobi_blocks(const double* A, int64_t n, int64_t lda, double* jac, hipStream_t st);
hipError_t spai_apply(const double* jac, const float* Y, const MeshFeat& f, double invG, const double* r, double* z,
                      hipStream_t st);
// E: n x S x 3 doubles of workspace
hipError_t spai_loss_grad(const double* jac, const float* Y, const MeshFeat& f, double invG, const double* B,
                          const double* X, int S, double* E, float* dY, double* lossp, hipStream_t st);
hipError_t randn(double* x, int64_t n, uint64_t seed, hipStream_t st);
hipError_t device_sum(const double* a, int64_t n, double* out, hipStream_t st);
void gemm_f32(bool ta, bool tb, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C,
              int ldc, const float* bias, int act, int accumulate, hipStream_t st);

}  // namespace cdj
