// C++17 drop-in adapter over librerdmft_gpu's C ABI (include/rerdmft_gpu.h).
//
// The reference selects its Fock builder by C++ overload resolution on the ERI-source type
// (nonRelFockMatrix(h, RiNonRelEriSource, P), NonRelFockMatrixRi.h:39-40; x2cFockMatrix(h,
// RiX2cEriSource, P), X2C_FockMatrixRi.h:38-40; rkbFockMatrix(h, RiDirectEriSource, P),
// UkbFockMatrixRi.h:96-98) and calls them from templated SCF loops
// (NonRelHartreeFock.cpp:132, X2C_HF.cpp:77, C4_DHF.cpp:81). This header supplies GPU source
// types with the same overload names, so those loops compile unchanged against them: construct the
// source once (the 3-index tensor is uploaded and stays resident in HBM), then every call is one
// synchronous GPU Fock build.
//
// The matrix type is a template parameter: any row-major dense matrix with rows(), cols(),
// data() and a (rows, cols) constructor -- the reference's Matrix<T> (src/Matrix.h:13-38) fits.
// Errors raise std::runtime_error with the library's message (the reference's shape-check texts).
#pragma once

#include <complex>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/rerdmft_gpu.h"

namespace rerdmft {
namespace gpu {

// One device context; shared by every source built on it.
class Device {
 public:
  explicit Device(int device = 0) {
    if (rr_create(device, &h_) != RR_OK) throw std::runtime_error("rr_create failed on device " + std::to_string(device));
  }
  ~Device() {
    if (h_) rr_destroy(h_);
  }
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;
  rr_handle handle() const { return h_; }
  void check(int rc) const {
    if (rc != RR_OK) throw std::runtime_error(rr_last_error(h_));
  }
  int allocate_slot() { return next_slot_++; }

 private:
  rr_handle h_ = nullptr;
  int next_slot_ = 0;
};

namespace detail {
template <class Mat>
const double* dptr(const Mat& m) {
  return reinterpret_cast<const double*>(m.data());
}
template <class Mat>
double* dptr(Mat& m) {
  return reinterpret_cast<double*>(m.data());
}
inline unsigned copies_for(int precision) {
  return RR_COPY_FP64 | (precision == RR_FP32 ? RR_COPY_FP32 : 0u) | (precision == RR_BF16 ? RR_COPY_BF16 : 0u);
}
}  // namespace detail

// ---- NON_REL: counterpart of RiNonRelEriSource (NonRelFockMatrixRi.h:30-37) ------------------
template <class RealMat>
struct GpuRiNonRelEriSource {
  std::shared_ptr<Device> dev;
  int slot = -1;
  std::size_t nl = 0;
  RealMat x_large;
  RealMat overlap;
  int precision = RR_FP64;

  GpuRiNonRelEriSource(std::shared_ptr<Device> d, std::size_t nl_, const RealMat& eri3_L, RealMat x, RealMat s,
                       int prec = RR_FP64)
      : dev(std::move(d)), nl(nl_), x_large(std::move(x)), overlap(std::move(s)), precision(prec) {
    slot = dev->allocate_slot();
    dev->check(rr_tensor_upload(dev->handle(), slot, detail::dptr(eri3_L), eri3_L.rows(), nl, detail::copies_for(prec)));
  }
};

template <class RealMat>
RealMat nonRelFockMatrix(const RealMat& h_core, const GpuRiNonRelEriSource<RealMat>& eri, const RealMat& density) {
  RealMat fock(h_core.rows(), h_core.cols());
  eri.dev->check(rr_nonrel_fock_ri(eri.dev->handle(), eri.slot, eri.precision, density.rows(), detail::dptr(h_core),
                                   detail::dptr(eri.x_large), eri.x_large.cols(), detail::dptr(eri.overlap),
                                   detail::dptr(density), detail::dptr(fock), nullptr, nullptr));
  return fock;
}

// ---- X2C: counterpart of RiX2cEriSource (X2C_FockMatrixRi.h:29-36) ----------------------------
template <class RealMat, class CplxMat>
struct GpuRiX2cEriSource {
  std::shared_ptr<Device> dev;
  int slot = -1;
  std::size_t nl = 0;
  CplxMat x_large;
  CplxMat overlap;
  int precision = RR_FP64;

  GpuRiX2cEriSource(std::shared_ptr<Device> d, std::size_t nl_, const RealMat& eri3_L, CplxMat x, CplxMat s,
                    int prec = RR_FP64)
      : dev(std::move(d)), nl(nl_), x_large(std::move(x)), overlap(std::move(s)), precision(prec) {
    slot = dev->allocate_slot();
    dev->check(rr_tensor_upload(dev->handle(), slot, detail::dptr(eri3_L), eri3_L.rows(), nl, detail::copies_for(prec)));
  }
};

template <class RealMat, class CplxMat>
CplxMat x2cFockMatrix(const CplxMat& h_x2c, const GpuRiX2cEriSource<RealMat, CplxMat>& eri, const CplxMat& density) {
  CplxMat fock(h_x2c.rows(), h_x2c.cols());
  eri.dev->check(rr_x2c_fock_ri(eri.dev->handle(), eri.slot, eri.precision, eri.nl, detail::dptr(h_x2c),
                                detail::dptr(eri.x_large), eri.x_large.cols(), detail::dptr(eri.overlap),
                                detail::dptr(density), detail::dptr(fock)));
  return fock;
}

// ---- 4c RKB: counterpart of RiDirectEriSource (UkbFockMatrixRi.h:84-89) -----------------------
template <class RealMat, class CplxMat>
struct GpuRiDirectEriSource {
  std::shared_ptr<Device> dev;
  int slot_ll = -1, slot_ss = -1;
  std::size_t nl = 0, ns = 0;
  CplxMat v_total, s_full, x_full;
  int precision = RR_FP64;

  GpuRiDirectEriSource(std::shared_ptr<Device> d, std::size_t nl_, std::size_t ns_, const RealMat& eri3_LL,
                       const RealMat& eri3_SS, CplxMat v, CplxMat s, CplxMat x, int prec = RR_FP64)
      : dev(std::move(d)), nl(nl_), ns(ns_), v_total(std::move(v)), s_full(std::move(s)), x_full(std::move(x)),
        precision(prec) {
    slot_ll = dev->allocate_slot();
    slot_ss = dev->allocate_slot();
    dev->check(rr_tensor_upload(dev->handle(), slot_ll, detail::dptr(eri3_LL), eri3_LL.rows(), nl, detail::copies_for(prec)));
    dev->check(rr_tensor_upload(dev->handle(), slot_ss, detail::dptr(eri3_SS), eri3_SS.rows(), ns, detail::copies_for(prec)));
  }
};

template <class RealMat, class CplxMat>
CplxMat rkbFockMatrix(const CplxMat& h_rkb, const GpuRiDirectEriSource<RealMat, CplxMat>& eri, const CplxMat& density) {
  CplxMat fock(h_rkb.rows(), h_rkb.cols());
  eri.dev->check(rr_rkb_fock_ri(eri.dev->handle(), eri.slot_ll, eri.slot_ss, eri.precision, eri.nl, eri.ns,
                                detail::dptr(h_rkb), detail::dptr(eri.v_total), eri.v_total.cols(), detail::dptr(eri.s_full),
                                detail::dptr(eri.x_full), eri.x_full.cols(), detail::dptr(density), detail::dptr(fock)));
  return fock;
}

// riFockTwoElectronDirect (UkbFockMatrixRi.h:70-75) on a resident source.
template <class RealMat, class CplxMat>
CplxMat riFockTwoElectronDirect(const GpuRiDirectEriSource<RealMat, CplxMat>& eri, const CplxMat& density_ukb,
                                const CplxMat& c_occ_ukb, const std::vector<double>& occupations) {
  CplxMat g(density_ukb.rows(), density_ukb.cols());
  eri.dev->check(rr_ri_fock_two_electron_direct(eri.dev->handle(), eri.slot_ll, eri.slot_ss, eri.precision, eri.nl, eri.ns,
                                                detail::dptr(density_ukb), detail::dptr(c_occ_ukb), occupations.size(),
                                                occupations.data(), detail::dptr(g)));
  return g;
}

// buildRiMoEri (RiMoEri.h:81-91): returns the (n_aux, n_total^2) complex MO tensor in host memory.
template <class RealMat, class CplxMat>
CplxMat buildRiMoEri(const GpuRiDirectEriSource<RealMat, CplxMat>& eri, const CplxMat& c_ukb, std::size_t n_aux) {
  const std::size_t nt = c_ukb.cols();
  CplxMat out(n_aux, nt * nt);
  eri.dev->check(rr_build_ri_mo_eri(eri.dev->handle(), eri.slot_ll, eri.slot_ss, eri.precision, RR_FP64, eri.nl, eri.ns,
                                    detail::dptr(c_ukb), nt, out.data()));
  return out;
}

// ---- DIIS: counterpart of Diis<T> and scfCommutatorError (Utils/DIIS.h:45-70), with the
// (error, Fock) history kept in device memory ------------------------------------------------------
namespace detail {
template <class Mat>
constexpr std::size_t scalar_doubles() {
  using T = typename std::remove_cv<typename std::remove_pointer<decltype(std::declval<const Mat&>().data())>::type>::type;
  return sizeof(T) / sizeof(double);  // 1 for double, 2 for std::complex<double>
}
}  // namespace detail

template <class Mat>
Mat scfCommutatorError(Device& dev, const Mat& fock, const Mat& density, const Mat& overlap) {
  Mat e(fock.rows(), fock.cols());
  dev.check(rr_scf_commutator_error(dev.handle(), detail::scalar_doubles<Mat>() == 2, fock.rows(), detail::dptr(fock),
                                    detail::dptr(density), detail::dptr(overlap), detail::dptr(e)));
  return e;
}

template <class Mat>
class GpuDiis {
 public:
  GpuDiis(std::shared_ptr<Device> dev, std::size_t max_vectors) : dev_(std::move(dev)), max_(max_vectors) {
    dev_->check(rr_diis_create(dev_->handle(), max_vectors, &d_));
  }
  ~GpuDiis() {
    if (d_) rr_diis_destroy(dev_->handle(), d_);
  }
  GpuDiis(const GpuDiis&) = delete;
  GpuDiis& operator=(const GpuDiis&) = delete;
  Mat extrapolate(const Mat& error, const Mat& value) {
    constexpr std::size_t k = detail::scalar_doubles<Mat>();
    Mat out(value.rows(), value.cols());
    double w[16];
    std::size_t nw = 0;
    dev_->check(rr_diis_extrapolate(dev_->handle(), d_, k * error.rows() * error.cols(), detail::dptr(error),
                                    k * value.rows() * value.cols(), detail::dptr(value), detail::dptr(out), w, &nw));
    last_.assign(w, w + nw);
    return out;
  }
  void clear() {
    dev_->check(rr_diis_clear(dev_->handle(), d_));
    last_.clear();
  }
  std::size_t maxVectors() const { return max_; }
  const std::vector<double>& lastWeights() const { return last_; }

 private:
  std::shared_ptr<Device> dev_;
  rr_diis_t d_ = nullptr;
  std::size_t max_;
  std::vector<double> last_;
};

}  // namespace gpu
}  // namespace rerdmft
