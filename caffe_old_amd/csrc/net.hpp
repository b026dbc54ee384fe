// net.hpp -- the learned multilevel pressure-correction preconditioner: network
// definition, device-resident network object and training set, and the launchers
// shared by net.hip (network kernels) and learned.hip (two-grid cycle, training loss,
// flexible CG).
#pragma once
#include <string>
#include <vector>

#include "internal.hpp"

namespace caffe {

// ---------------------------------------------------------------- definition
// Twelve 3x3x3 convolutions (zero padding, no bias, identity activation: the map is
// linear, as a Krylov preconditioner must be).  Level 0 is the coarse grid (half the
// fine grid in each direction); levels 1..3 halve it again by 2x2x2 averaging; the
// up path adds nearest-neighbour up-sampled coarse features to the skip features.
enum { MODE_PLAIN = 0, MODE_POOL = 1, MODE_UPADD = 2 };
struct LayerDef {
    int cin, cout, level, mode;
    int src;    // layer whose output feeds this one (-1: the network input)
    int skip;   // UPADD: layer whose output is added after up-sampling src (-1: none)
};
constexpr int NET_LAYERS = 12;
constexpr int NET_IN = 0, NET_OUT = 11;
constexpr LayerDef kNet[NET_LAYERS] = {
    {1, 16, 0, MODE_PLAIN, -1, -1},    // in
    {16, 16, 0, MODE_PLAIN, 0, -1},    // d0
    {16, 32, 1, MODE_POOL, 1, -1},     // d1a
    {32, 32, 1, MODE_PLAIN, 2, -1},    // d1b
    {32, 64, 2, MODE_POOL, 3, -1},     // d2a
    {64, 64, 2, MODE_PLAIN, 4, -1},    // d2b
    {64, 64, 3, MODE_POOL, 5, -1},     // d3a
    {64, 64, 3, MODE_PLAIN, 6, -1},    // d3b
    {64, 32, 2, MODE_UPADD, 7, 5},     // u2
    {32, 16, 1, MODE_UPADD, 8, 3},     // u1
    {16, 16, 0, MODE_UPADD, 9, 1},     // u0
    {16, 1, 0, MODE_PLAIN, 10, -1},    // out
};
constexpr int NET_ZERO_SLOT = 63;  // the forward pass's "not done" flag when the caller has none
constexpr float NET_OMEGA = 0.8f;  // damped-Jacobi weight of the fine-grid smoothing sweeps

// One training pair: the p' system (FP32 symmetric 4-array storage on the padded fine
// grid), its right-hand side scaled to unit norm, the converged correction with the
// same scale, and x^T A x.
struct Sample {
    Dims D;
    float* A = nullptr;
    float* b = nullptr;
    float* x = nullptr;
    double nrm = 0.;
};

struct Net {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;

    long nParams = 0;
    long off[NET_LAYERS] = {};       // FP32 parameter offset of each layer, [27][cout][cin]
    long poff[NET_LAYERS] = {};      // packed BF16 offset of each MFMA layer
    float* theta = nullptr;          // FP32 master parameters
    float* grad = nullptr;
    float* m1 = nullptr;             // Adam moments
    float* m2 = nullptr;
    uint16_t* wf = nullptr;          // BF16, forward packing
    uint16_t* wb = nullptr;          // BF16, input-gradient packing (flipped taps, swapped channels)
    float* wpart = nullptr;          // per-chunk weight-gradient partial sums
    long wpartCap = 0;
    int adamT = 0;
    int post = 1;                    // post-smoothing sweeps of the cycle in training, loss and precond
    bool packed = false;

    // activations for the current coarse dimensions
    int nx = 0, ny = 0, nz = 0;
    bool trainBufs = false;
    uint16_t* act[NET_LAYERS] = {};  // BF16 layer outputs (channels last); NET_OUT unused
    uint16_t* pooled[NET_LAYERS] = {};  // 2x2x2 averages of a layer output that a POOL layer reads (BF16)
    float* gact[NET_LAYERS] = {};    // FP32 gradients of the layer outputs
    float* gstage = nullptr;         // FP32 gradient of a layer's staged input
    float* vin = nullptr;            // coarse input (FP32)
    float* vout = nullptr;           // coarse output (FP32)
    float* gout = nullptr;           // gradient of the coarse output
    std::vector<void*> dimAllocs;

    // fine-grid work for the training loss and the flexible CG
    int fineP = 0;
    float* fw = nullptr;             // 12 fine vectors
    std::vector<void*> fineAllocs;

    // reductions
    double* part = nullptr;          // 2 x MAX_RED_BLOCKS x RED_STRIDE
    double* scal = nullptr;          // 64 device doubles; slot NET_ZERO_SLOT stays zero
    double* hscal = nullptr;         // pinned mirror
    float* losses = nullptr;         // device loss history
    long lossCap = 0;

    std::vector<Sample> data;
    std::vector<void*> allocs;
};

// ---- network (net.hip)
void net_set_dims(Net& n, int nx, int ny, int nz, bool train);
void net_pack(Net& n, hipStream_t s);
// out = net(vin) on the coarse grid set by net_set_dims; kernels return at once when
// *done != 0 (done may be null)
void net_forward(Net& n, hipStream_t s, const float* vin, float* vout, const double* done);
// grad += d<gout, net(vin)>/dtheta (the forward pass must have run on vin just before)
void net_backward(Net& n, hipStream_t s, const float* vin, const float* gout);
void net_adam(Net& n, hipStream_t s, float lr, float gscale);
void* net_alloc(Net& n, size_t bytes, std::vector<void*>& list);

// ---- two-grid cycle, loss and training (learned.hip)
void tg_ensure_fine(Net& n, const Dims& D);
// z = learned two-grid cycle applied to r (fine FP32 padded vectors)
void tg_apply(Net& n, hipStream_t s, const Dims& D, const float* A, const float* r, float* z, const double* done);
// loss of one sample; grad += dL/dtheta when withGrad; the loss lands in *lossDev
void tg_loss(Net& n, hipStream_t s, const Sample& smp, bool withGrad, float* lossDev);
int learned_fcg(Ctx& c, double shift, double bNorm, double rtol, int maxit, double delta, int* refs, double* relOut,
                bool* meanQueued);

}  // namespace caffe

struct caffe_net {
    caffe::Net n;
};
