// Learned coarse basis of the batched two-level preconditioner: a small graph
// network on the element-edge graph of the mesh, trained on the GPU on
// (conductivity field, solution) pairs produced by the solver itself.  There
// is no counterpart in the reference.
//
//     He0 = relu(Xe We0 + be0)            Hf0 = relu(Xf Wf0 + bf0)
//     layer l = 1, 2, 3:
//         Mf  = mean over the elements of an edge of He(l-1)
//         Hf(l) = relu(Hf(l-1) A1_l + Mf A2_l + a_l)
//         l < 3:  Me = mean over the edges of an element of Hf(l)
//                 He(l) = relu(He(l-1) B1_l + Me B2_l + b_l)
//     phi = Hf3 w_out + c_out            (the second coarse column per edge)
//
// Hidden width 64.  Activations and the GEMM operands are BF16; every product
// runs on the MFMA units (v_mfma_f32_16x16x16_bf16) with FP32 accumulation;
// master weights, gradients and the Adam moments are FP32.  The training loss
// is the part of the solution lambda that the coarse columns {1, phi} of each
// aggregate cannot represent (least squares per aggregate, normalised).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "batched.hh"

namespace fl {

class GnnGPU {
   public:
    static constexpr int kH = 64, kNF = 16;
    GnnGPU(DarcyBatchGPU &batch, unsigned seed);
    ~GnnGPU();
    int n_params() const { return np_; }
    void get_params(float *out) const;  // host, flat FP32 in the documented order
    void set_params(const float *in);
    // one realization: K [E] (device, FP64) -> phi [F] (device, FP32, edge order)
    void forward(const double *dK, float *dphi_out);
    // loss of the current forward against lambda [F] (device FP64, edge order),
    // gradients accumulated into the FP32 gradient vector; returns the loss
    double loss_backward(const double *dlambda);
    void zero_grad();
    void get_grad(float *out) const;  // host
    void adam_step(double lr);
    // features of the last forward (host copies, FP32 [E][16], [F][16])
    void get_features(float *xe, float *xf) const;
    // inference for R realizations: K [R][E] -> phi [R][F] (device)
    void infer(const double *dK, int R, float *dphi);
    int n_edges() const { return F_; }
    int n_elements() const { return E_; }

   private:
    void pack_weights();
    void gemm_fwd(int N, const unsigned short *X1, int K1, int w1, const unsigned short *X2, int w2, int bias,
                  unsigned short *Y);
    DarcyBatchGPU &b_;
    hipStream_t stream_ = nullptr;
    int E_, F_, Ep_, Fp_, nagg_, np_;
    std::vector<void *> owned_;
    std::vector<int> off_;  // parameter offsets in the flat vector
    // graph
    int *d_eptr_ = nullptr, *d_eedge_ = nullptr, *d_fptr_ = nullptr, *d_fel_ = nullptr;
    int *d_aptr_ = nullptr, *d_aedge_ = nullptr, *d_dim_ = nullptr;
    float *d_Xf_ = nullptr;  // FP32 edge features (K independent) before BF16 packing
    unsigned short *d_xe_ = nullptr, *d_xf_ = nullptr;
    // parameters
    float *d_P_ = nullptr, *d_G_ = nullptr, *d_m_ = nullptr, *d_v_ = nullptr;
    unsigned short *d_W_ = nullptr, *d_Wt_ = nullptr;  // BF16 copies [K][64] and [64][K] per matrix
    // activations (BF16)
    unsigned short *d_He_[3] = {}, *d_Hf_[4] = {}, *d_Mf_[4] = {}, *d_Me_[3] = {};
    float *d_phi_ = nullptr;
    // backward work (FP32 except the BF16 dZ)
    float *d_dHf_[2] = {}, *d_dHe_[2] = {}, *d_dM_ = nullptr, *d_wpart_ = nullptr, *d_lossbuf_ = nullptr;
    unsigned short *d_dZ_ = nullptr;
    double *d_stats_ = nullptr;
    double *d_lkpart_ = nullptr;  // partial ln K sums
    int adam_t_ = 0;
};

}  // namespace fl
