// pybind11 host module `sapien_amd._pinocchio`: the drop-in PinocchioModel
// (same class, method and argument names, defaults, return types and error
// types as the reference binding, pinocchio/pinocchio_model.cpp:306-367) and a
// FP32 sapien.Pose stand-in (include/sapien/math/pose.h:24-49), plus batched
// methods over (B, dof) tensors consumed through DLPack. Every computation goes
// through the C ABI of include/sapien_amd.h, which runs HIP kernels on gfx950.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <array>
#include <cmath>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "dlpack_abi.h"
#include "sapien_amd.h"

namespace py = pybind11;

namespace {

void check(int rc) {
  if (rc == SAPAMD_OK) return;
  std::string msg = sapamd_last_error();
  if (rc == SAPAMD_INVALID_ARGUMENT) throw py::value_error(msg);
  throw std::runtime_error(msg);
}

using DArr = py::array_t<double, py::array::c_style | py::array::forcecast>;
using IArr = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;

// ------------------------------------------------------------------ Pose (FP32, q = wxyz)
struct Pose {
  std::array<float, 3> p{0, 0, 0};
  std::array<float, 4> q{1, 0, 0, 0};

  std::array<float, 3> rotate(const std::array<float, 3> &v) const {
    // 2 (u.v) u + (w^2 - u.u) v + 2 w (u x v)  (include/sapien/math/quat.h:76-79)
    float w = q[0], x = q[1], y = q[2], z = q[3];
    float uv = x * v[0] + y * v[1] + z * v[2], uu = x * x + y * y + z * z;
    std::array<float, 3> c = {y * v[2] - z * v[1], z * v[0] - x * v[2], x * v[1] - y * v[0]};
    return {2 * uv * x + (w * w - uu) * v[0] + 2 * w * c[0], 2 * uv * y + (w * w - uu) * v[1] + 2 * w * c[1],
            2 * uv * z + (w * w - uu) * v[2] + 2 * w * c[2]};
  }
  Pose mul(const Pose &o) const {
    Pose r;
    auto rp = rotate(o.p);
    for (int i = 0; i < 3; ++i) r.p[i] = rp[i] + p[i];
    const auto &a = q, &b = o.q;
    r.q = {a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3], a[0] * b[1] + b[0] * a[1] + a[2] * b[3] - b[2] * a[3],
           a[0] * b[2] + b[0] * a[2] + a[3] * b[1] - b[3] * a[1], a[0] * b[3] + b[0] * a[3] + a[1] * b[2] - b[1] * a[2]};
    return r;
  }
  Pose inv() const {
    Pose r;
    r.q = {q[0], -q[1], -q[2], -q[3]};
    auto np = r.rotate({-p[0], -p[1], -p[2]});
    r.p = np;
    return r;
  }
  std::array<double, 9> matrix() const {
    double n = std::sqrt((double)q[0] * q[0] + (double)q[1] * q[1] + (double)q[2] * q[2] + (double)q[3] * q[3]);
    double w = q[0] / n, x = q[1] / n, y = q[2] / n, z = q[3] / n;
    return {1 - 2 * (y * y + z * z), 2 * (x * y - w * z),     2 * (x * z + w * y),
            2 * (x * y + w * z),     1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
            2 * (x * z - w * y),     2 * (y * z + w * x),     1 - 2 * (x * x + y * y)};
  }
  static std::array<float, 4> quat_of(const double *R) {
    // Shoemake's method (largest of w, x, y, z first), normalized
    double t = R[0] + R[4] + R[8], o[4];
    if (t > 0) {
      double r = std::sqrt(t + 1), h = 0.5 / r;
      o[0] = 0.5 * r; o[1] = (R[7] - R[5]) * h; o[2] = (R[2] - R[6]) * h; o[3] = (R[3] - R[1]) * h;
    } else {
      int i = 0;
      if (R[4] > R[0]) i = 1;
      if (R[8] > R[3 * i + i]) i = 2;
      int j = (i + 1) % 3, k = (j + 1) % 3;
      double r = std::sqrt(R[3 * i + i] - R[3 * j + j] - R[3 * k + k] + 1), h = 0.5 / r;
      o[1 + i] = 0.5 * r;
      o[1 + j] = (R[3 * j + i] + R[3 * i + j]) * h;
      o[1 + k] = (R[3 * k + i] + R[3 * i + k]) * h;
      o[0] = (R[3 * k + j] - R[3 * j + k]) * h;
    }
    double n = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2] + o[3] * o[3]);
    return {float(o[0] / n), float(o[1] / n), float(o[2] / n), float(o[3] / n)};
  }
  std::array<float, 3> rpy() const {
    auto R = matrix();  // R = Rz(y) Ry(p) Rx(r)
    double pitch = std::asin(std::max(-1.0, std::min(1.0, -R[6])));
    double roll, yaw;
    if (std::abs(R[6]) < 1 - 1e-9) {
      roll = std::atan2(R[7], R[8]);
      yaw = std::atan2(R[3], R[0]);
    } else {
      roll = 0;
      yaw = std::atan2(-R[1], R[4]);
    }
    return {float(roll), float(pitch), float(yaw)};
  }
  void set_rpy(const std::array<float, 3> &v) {
    double cr = std::cos(v[0] / 2.0), sr = std::sin(v[0] / 2.0), cp = std::cos(v[1] / 2.0), sp = std::sin(v[1] / 2.0),
           cy = std::cos(v[2] / 2.0), sy = std::sin(v[2] / 2.0);
    // q = qz(y) * qy(p) * qx(r)  (include/sapien/math/conversion.h:59-61)
    q = {float(cy * cp * cr + sy * sp * sr), float(cy * cp * sr - sy * sp * cr), float(cy * sp * cr + sy * cp * sr),
         float(sy * cp * cr - cy * sp * sr)};
  }
};

Pose pose_from_object(const py::object &o) {
  if (py::isinstance<Pose>(o)) return o.cast<Pose>();
  Pose r;
  auto p = o.attr("p").cast<std::vector<float>>();
  auto q = o.attr("q").cast<std::vector<float>>();
  if (p.size() != 3 || q.size() != 4) throw py::value_error("pose needs p (3) and q (4)");
  for (int i = 0; i < 3; ++i) r.p[i] = p[i];
  for (int i = 0; i < 4; ++i) r.q[i] = q[i];
  return r;
}

// ------------------------------------------------------------------ DLPack intake
struct DLIn {
  dlabi::ManagedTensor *mt = nullptr;
  void *ptr = nullptr;
  std::vector<int64_t> strides;  // elements, filled by take_strided
  ~DLIn() {
    if (mt && mt->deleter) mt->deleter(mt);
  }
};

std::unique_ptr<DLIn> take(const py::object &t, uint8_t code, const std::vector<int64_t> &shape, int device) {
  py::object cap = t.attr("__dlpack__")();
  PyObject *c = cap.ptr();
  if (!PyCapsule_IsValid(c, "dltensor")) throw py::type_error("__dlpack__ did not return a dltensor capsule");
  auto *mt = static_cast<dlabi::ManagedTensor *>(PyCapsule_GetPointer(c, "dltensor"));
  PyCapsule_SetName(c, "used_dltensor");
  auto in = std::make_unique<DLIn>();
  in->mt = mt;
  const dlabi::Tensor &d = mt->dl_tensor;
  if (d.device.device_type != dlabi::kROCM)
    throw py::value_error("batched methods need a GPU tensor (DLPack device kDLROCM)");
  if (d.device.device_id != device) throw py::value_error("tensor is on a different GPU than the model");
  if (d.dtype.code != code || d.dtype.bits != 32 || d.dtype.lanes != 1) throw py::value_error("unexpected dtype");
  if (d.ndim != (int32_t)shape.size()) throw py::value_error("unexpected tensor rank");
  int64_t numel = 1;
  for (int i = 0; i < d.ndim; ++i) numel *= d.shape[i];
  int64_t stride = 1;
  for (int i = d.ndim - 1; i >= 0; --i) {
    if (shape[i] >= 0 && d.shape[i] != shape[i]) throw py::value_error("unexpected tensor shape");
    if (numel > 0 && d.strides && d.shape[i] > 1 && d.strides[i] != stride)
      throw py::value_error("tensor must be contiguous");
    stride *= d.shape[i];
  }
  in->ptr = static_cast<char *>(d.data) + d.byte_offset;
  return in;
}

// A float32 GPU tensor whose last dimension is contiguous and whose other
// dimensions may be strided (padded buffers, slices of them): shape entries
// < 0 are free; returns the element strides.
std::unique_ptr<DLIn> take_strided(const py::object &t, const std::vector<int64_t> &shape, int device) {
  py::object cap = t.attr("__dlpack__")();
  PyObject *c = cap.ptr();
  if (!PyCapsule_IsValid(c, "dltensor")) throw py::type_error("__dlpack__ did not return a dltensor capsule");
  auto *mt = static_cast<dlabi::ManagedTensor *>(PyCapsule_GetPointer(c, "dltensor"));
  PyCapsule_SetName(c, "used_dltensor");
  auto in = std::make_unique<DLIn>();
  in->mt = mt;
  const dlabi::Tensor &d = mt->dl_tensor;
  if (d.device.device_type != dlabi::kROCM)
    throw py::value_error("batched methods need a GPU tensor (DLPack device kDLROCM)");
  if (d.device.device_id != device) throw py::value_error("tensor is on a different GPU than the model");
  if (d.dtype.code != dlabi::kFloat || d.dtype.bits != 32 || d.dtype.lanes != 1)
    throw py::value_error("strided buffers must be float32");
  if (d.ndim != (int32_t)shape.size()) throw py::value_error("unexpected tensor rank");
  in->strides.assign(d.ndim, 0);
  int64_t dense = 1;
  for (int i = d.ndim - 1; i >= 0; --i) {
    if (shape[i] >= 0 && d.shape[i] != shape[i]) throw py::value_error("unexpected tensor shape");
    in->strides[i] = d.strides ? d.strides[i] : dense;
    dense *= d.shape[i];
  }
  if (d.ndim > 0 && d.shape[d.ndim - 1] > 1 && in->strides[d.ndim - 1] != 1)
    throw py::value_error("the last dimension of a strided buffer must be contiguous");
  for (int i = 0; i < d.ndim; ++i)
    if (in->strides[i] < 0) throw py::value_error("negative strides are not supported");
  in->ptr = static_cast<char *>(d.data) + d.byte_offset;
  return in;
}

// ------------------------------------------------------------------ model
class PyModel {
 public:
  PyModel(const std::string &urdf, const std::vector<double> &gravity) {
    if (gravity.size() != 3) throw py::value_error("gravity must have 3 entries");
    check(sapamd_model_create(urdf.c_str(), gravity.data(), &m_));
  }
  ~PyModel() { sapamd_model_destroy(m_); }
  PyModel(const PyModel &) = delete;

  sapamd_info info() const {
    sapamd_info i;
    check(sapamd_model_info(m_, &i));
    return i;
  }
  int nv() const { return info().nv; }

  void set_link_order(const std::vector<std::string> &names) {
    std::vector<const char *> p;
    for (auto &s : names) p.push_back(s.c_str());
    check(sapamd_set_link_order(m_, p.data(), (int32_t)p.size()));
  }
  void set_joint_order(const std::vector<std::string> &names) {
    std::vector<const char *> p;
    for (auto &s : names) p.push_back(s.c_str());
    check(sapamd_set_joint_order(m_, p.data(), (int32_t)p.size()));
  }
  void compute_forward_kinematics(const DArr &q) { check(sapamd_compute_forward_kinematics(m_, q.data(), q.size())); }
  Pose get_link_pose(int64_t link) {
    float p7[7];
    check(sapamd_get_link_pose(m_, (int32_t)clamp_index(link), p7));
    Pose r;
    for (int i = 0; i < 3; ++i) r.p[i] = p7[i];
    for (int i = 0; i < 4; ++i) r.q[i] = p7[3 + i];
    return r;
  }
  void compute_full_jacobian(const DArr &q) { check(sapamd_compute_full_jacobian(m_, q.data(), q.size())); }
  py::array_t<double> get_link_jacobian(int64_t link, bool local) {
    int n = nv();
    py::array_t<double> out({6, n});
    check(sapamd_get_link_jacobian(m_, (int32_t)clamp_index(link), local, out.mutable_data()));
    return out;
  }
  py::array_t<double> compute_single_link_local_jacobian(const DArr &q, int64_t link) {
    int n = nv();
    py::array_t<double> out({6, n});
    check(sapamd_compute_single_link_local_jacobian(m_, q.data(), q.size(), (int32_t)clamp_index(link),
                                                    out.mutable_data()));
    return out;
  }
  py::array_t<double> compute_generalized_mass_matrix(const DArr &q) {
    int n = nv();
    py::array_t<double> out({n, n});
    check(sapamd_compute_generalized_mass_matrix(m_, q.data(), q.size(), out.mutable_data()));
    return out;
  }
  py::array_t<double> compute_coriolis_matrix(const DArr &q, const DArr &qd) {
    int n = nv();
    same_len(q, qd);
    py::array_t<double> out({n, n});
    check(sapamd_compute_coriolis_matrix(m_, q.data(), qd.data(), q.size(), out.mutable_data()));
    return out;
  }
  py::array_t<double> compute_inverse_dynamics(const DArr &q, const DArr &qd, const DArr &qdd) {
    same_len(q, qd);
    same_len(q, qdd);
    py::array_t<double> out(nv());
    check(sapamd_compute_inverse_dynamics(m_, q.data(), qd.data(), qdd.data(), q.size(), out.mutable_data()));
    return out;
  }
  py::array_t<double> compute_forward_dynamics(const DArr &q, const DArr &qd, const DArr &qf) {
    same_len(q, qd);
    same_len(q, qf);
    py::array_t<double> out(nv());
    check(sapamd_compute_forward_dynamics(m_, q.data(), qd.data(), qf.data(), q.size(), out.mutable_data()));
    return out;
  }
  py::tuple compute_inverse_kinematics(int64_t link, const py::object &pose, const DArr &q0, const IArr &mask,
                                       double eps, int max_iterations, double dt, double damp) {
    Pose P = pose_from_object(pose);
    float p7[7] = {P.p[0], P.p[1], P.p[2], P.q[0], P.q[1], P.q[2], P.q[3]};
    int n = nv();
    py::array_t<double> q(n), err(6);
    int32_t ok = 0;
    check(sapamd_compute_inverse_kinematics(m_, (int32_t)clamp_index(link), p7, q0.size() ? q0.data() : nullptr,
                                            q0.size(), mask.size() ? mask.data() : nullptr, mask.size(), eps,
                                            max_iterations, dt, damp, q.mutable_data(), &ok, err.mutable_data()));
    return py::make_tuple(q, bool(ok), err);
  }
  // A fresh seed per call from the model's own generator (seeded from the
  // system's entropy, as the reference's unseeded randomConfiguration draws),
  // the values drawn on the GPU.
  py::array_t<double> get_random_qpos() {
    py::array_t<double> out(nv());
    check(sapamd_get_random_qpos(m_, rng_(), out.mutable_data()));
    return out;
  }
  py::object random_qpos_batched(int64_t B, py::object seed, py::object stream) {
    StreamScope scope_(stream, info().device);
    if (B < 0) throw py::value_error("negative batch size");
    py::module_ torch = py::module_::import("torch");
    sapamd_info i = info();
    py::object out = torch.attr("empty")(py::make_tuple(B, i.nv), py::arg("device") = torch.attr("device")("cuda", i.device),
                                         py::arg("dtype") = torch.attr("float32"));
    const uint64_t sd = seed.is_none() ? rng_() : seed.cast<uint64_t>();
    auto vo = take(out, dlabi::kFloat, {B, i.nv}, i.device);
    void *s = stream.is_none() ? current_stream(i.device) : reinterpret_cast<void *>(stream.cast<intptr_t>());
    check(sapamd_random_qpos_batched(m_, B, sd, static_cast<float *>(vo->ptr), s));
    return out;
  }
  py::tuple joint_limits() {
    int n = nv();
    py::array_t<double> lo(n), hi(n);
    check(sapamd_joint_limits(m_, lo.mutable_data(), hi.mutable_data()));
    return py::make_tuple(lo, hi);
  }

  // ---------------------------------------------------------------- batched (DLPack, FP32)
  py::object fk_jac_rnea_batched(py::object q, py::object qd, py::object qdd, bool poses, bool jac, bool rnea,
                                 py::object stream) {
    StreamScope scope_(stream, info().device);
    Batch b(*this, q);
    py::object qdT, qddT, P = py::none(), J = py::none(), T = py::none();
    std::unique_ptr<DLIn> vqd, vqdd, vP, vJ, vT;
    if (rnea) {
      qdT = b.input(qd);
      qddT = b.input(qdd);
      vqd = b.view(qdT);
      vqdd = b.view(qddT);
      T = b.empty({b.B, b.n});
      vT = b.view(T);
    }
    if (poses) {
      P = b.empty({b.B, info().nlinks, 7});
      vP = b.view(P);
    }
    if (jac) {
      J = b.empty({b.B, 6, b.n});
      vJ = b.view(J);
    }
    auto vq = b.view(b.q);
    check(sapamd_fk_jac_rnea_batched(m_, b.B, f(vq), rnea ? f(vqd) : nullptr, rnea ? f(vqdd) : nullptr,
                                     poses ? f(vP) : nullptr, jac ? f(vJ) : nullptr, rnea ? f(vT) : nullptr,
                                     b.stream(stream)));
    return py::make_tuple(P, J, T);
  }
  py::object fk_batched(py::object q, py::object stream) {
    return fk_jac_rnea_batched(q, py::none(), py::none(), true, false, false, stream).cast<py::tuple>()[0];
  }
  py::object jac_batched(py::object q, py::object stream) {
    return fk_jac_rnea_batched(q, py::none(), py::none(), false, true, false, stream).cast<py::tuple>()[1];
  }
  py::object rnea_batched(py::object q, py::object qd, py::object qdd, py::object stream) {
    return fk_jac_rnea_batched(q, qd, qdd, false, false, true, stream).cast<py::tuple>()[2];
  }
  py::object link_jacobians_batched(py::object q, py::object stream) {
    StreamScope scope_(stream, info().device);
    Batch b(*this, q);
    int L = info().nlinks;
    py::object Jw = b.empty({b.B, L, 6, b.n}), Jl = b.empty({b.B, L, 6, b.n});
    auto vq = b.view(b.q);
    auto vw = b.view(Jw), vl = b.view(Jl);
    check(sapamd_link_jacobians_batched(m_, b.B, f(vq), f(vw), f(vl), b.stream(stream)));
    return py::make_tuple(Jw, Jl);
  }
  py::object crba_batched(py::object q, py::object stream) {
    StreamScope scope_(stream, info().device);
    Batch b(*this, q);
    py::object M = b.empty({b.B, b.n, b.n});
    auto vq = b.view(b.q);
    auto vM = b.view(M);
    check(sapamd_crba_batched(m_, b.B, f(vq), f(vM), b.stream(stream)));
    return M;
  }
  py::object coriolis_batched(py::object q, py::object qd, py::object stream) {
    StreamScope scope_(stream, info().device);
    Batch b(*this, q);
    py::object qdT = b.input(qd), C = b.empty({b.B, b.n, b.n});
    auto vq = b.view(b.q), vqd = b.view(qdT), vC = b.view(C);
    check(sapamd_coriolis_batched(m_, b.B, f(vq), f(vqd), f(vC), b.stream(stream)));
    return C;
  }
  py::object aba_batched(py::object q, py::object qd, py::object tau, py::object stream) {
    StreamScope scope_(stream, info().device);
    Batch b(*this, q);
    py::object qdT = b.input(qd), tT = b.input(tau), A = b.empty({b.B, b.n});
    auto vq = b.view(b.q), vqd = b.view(qdT), vt = b.view(tT), vA = b.view(A);
    check(sapamd_aba_batched(m_, b.B, f(vq), f(vqd), f(vt), f(vA), b.stream(stream)));
    return A;
  }
  py::object clik_batched(int64_t link, py::object target, py::object q0, py::object mask, double eps,
                          int max_iterations, double dt, double damp, py::object stream) {
    StreamScope scope_(stream, info().device);
    Batch b(*this, target, 7);
    py::object q0T = q0.is_none() ? py::none() : b.input(q0);
    py::object Q = b.empty({b.B, b.n}), E = b.empty({b.B, 6});
    py::object S = b.empty_i32({b.B}), I = b.empty_i32({b.B});
    py::object maskT = py::none();
    std::unique_ptr<DLIn> vmask, vq0;
    if (!mask.is_none()) {
      maskT = b.input_i32(mask, b.n);
      vmask = take(maskT, dlabi::kInt, {b.n}, b.device);
    }
    if (!q0T.is_none()) vq0 = b.view(q0T);
    auto vt = b.view(b.q), vQ = b.view(Q), vE = b.view(E);
    auto vS = take(S, dlabi::kInt, {b.B}, b.device), vI = take(I, dlabi::kInt, {b.B}, b.device);
    check(sapamd_clik_batched(m_, b.B, (int32_t)clamp_index(link), f(vt), vq0 ? f(vq0) : nullptr,
                              vmask ? static_cast<int32_t *>(vmask->ptr) : nullptr, eps, max_iterations, dt, damp,
                              f(vQ), static_cast<int32_t *>(vS->ptr), f(vE), static_cast<int32_t *>(vI->ptr),
                              b.stream(stream)));
    return py::make_tuple(Q, S.attr("bool")(), E, I);
  }

  // ---------------------------------------------------------------- padded buffers, passive forces
  // Writes into caller buffers (no allocation, no copies): poses (B, >= L, >= 7)
  // with the last dimension contiguous, e.g. the first 7 columns of a
  // [n_art, max_links, 13] body-state buffer; joint arrays (B, nv) with any row
  // stride, e.g. [:, :nv] of a [n_art, max_dof] buffer.
  void fk_jac_rnea_into(py::object q, py::object qd, py::object qdd, py::object poses, py::object J, py::object tau,
                        py::object stream) {
    StreamScope scope_(stream, info().device);
    sapamd_info i = info();
    auto vq = take_strided(q, {-1, i.nv}, i.device);
    const int64_t B = q.attr("shape").cast<std::vector<int64_t>>()[0];
    std::unique_ptr<DLIn> vqd, vqdd, vP, vJ, vT;
    if (!tau.is_none()) {
      vT = take_strided(tau, {B, i.nv}, i.device);
      vqd = take_strided(qd, {B, i.nv}, i.device);
      vqdd = take_strided(qdd, {B, i.nv}, i.device);
    }
    if (!poses.is_none()) {
      vP = take_strided(poses, {B, -1, -1}, i.device);
      auto shp = poses.attr("shape").cast<std::vector<int64_t>>();
      if (shp[1] < i.nlinks || shp[2] < 7) throw py::value_error("poses must have shape (B, >= nlinks, >= 7)");
    }
    if (!J.is_none()) vJ = take(J, dlabi::kFloat, {B, 6, i.nv}, i.device);
    void *s = stream.is_none() ? current_stream(i.device) : reinterpret_cast<void *>(stream.cast<intptr_t>());
    auto fp = [](const std::unique_ptr<DLIn> &v) { return v ? static_cast<float *>(v->ptr) : nullptr; };
    check(sapamd_fk_jac_rnea_strided(m_, B, fp(vq), vq->strides[0], fp(vqd), vqd ? vqd->strides[0] : 0, fp(vqdd),
                                     vqdd ? vqdd->strides[0] : 0, fp(vP), vP ? vP->strides[0] : 0,
                                     vP ? vP->strides[1] : 0, fp(vJ), fp(vT), vT ? vT->strides[0] : 0, s));
  }
  void link_state_into(py::object q, py::object qd, py::object body, py::object stream) {
    StreamScope scope_(stream, info().device);
    sapamd_info i = info();
    auto vq = take_strided(q, {-1, i.nv}, i.device);
    const int64_t B = q.attr("shape").cast<std::vector<int64_t>>()[0];
    std::unique_ptr<DLIn> vqd;
    if (!qd.is_none()) vqd = take_strided(qd, {B, i.nv}, i.device);
    auto vb = take_strided(body, {B, -1, -1}, i.device);
    auto shp = body.attr("shape").cast<std::vector<int64_t>>();
    if (shp[1] < i.nlinks || shp[2] < 13) throw py::value_error("body must have shape (B, >= nlinks, >= 13)");
    void *s = stream.is_none() ? current_stream(i.device) : reinterpret_cast<void *>(stream.cast<intptr_t>());
    check(sapamd_link_state_strided(m_, B, static_cast<float *>(vq->ptr), vq->strides[0],
                                    vqd ? static_cast<float *>(vqd->ptr) : nullptr, vqd ? vqd->strides[0] : 0,
                                    static_cast<float *>(vb->ptr), vb->strides[0], vb->strides[1], s));
  }
  py::object passive_force_batched(py::object q, py::object qd, bool gravity, bool coriolis, py::object out,
                                   py::object stream) {
    StreamScope scope_(stream, info().device);
    sapamd_info i = info();
    py::module_ torch = py::module_::import("torch");
    py::object dev = torch.attr("device")("cuda", i.device);
    auto as_gpu = [&](py::object x) {
      if (!py::isinstance(x, torch.attr("Tensor"))) x = torch.attr("as_tensor")(x);
      x = x.attr("detach")().attr("to")(py::arg("device") = dev, py::arg("dtype") = torch.attr("float32"));
      if (x.attr("dim")().cast<int>() == 1) x = x.attr("unsqueeze")(0);
      return x;
    };
    q = as_gpu(q);
    auto vq = take_strided(q, {-1, i.nv}, i.device);
    const int64_t B = q.attr("shape").cast<std::vector<int64_t>>()[0];
    std::unique_ptr<DLIn> vqd;
    if (coriolis) {
      if (qd.is_none()) throw py::value_error("qvel is needed for Coriolis and centrifugal forces");
      qd = as_gpu(qd);
      vqd = take_strided(qd, {B, i.nv}, i.device);
    }
    if (out.is_none())
      out = torch.attr("empty")(py::make_tuple(B, i.nv), py::arg("device") = dev, py::arg("dtype") = torch.attr("float32"));
    auto vo = take_strided(out, {B, i.nv}, i.device);
    void *s = stream.is_none() ? current_stream(i.device) : reinterpret_cast<void *>(stream.cast<intptr_t>());
    check(sapamd_passive_force_batched(m_, B, static_cast<float *>(vq->ptr), vq->strides[0],
                                       vqd ? static_cast<float *>(vqd->ptr) : nullptr, vqd ? vqd->strides[0] : 0,
                                       gravity ? 1 : 0, coriolis ? 1 : 0, static_cast<float *>(vo->ptr),
                                       vo->strides[0], s));
    return out;
  }
  static void *current_stream(int device) {
    py::module_ torch = py::module_::import("torch");
    py::object cs = torch.attr("cuda").attr("current_stream")(torch.attr("device")("cuda", device));
    return reinterpret_cast<void *>(cs.attr("cuda_stream").cast<intptr_t>());
  }

  // Multi-start CLIK, raw: q0 (nstart * B, nv) start-major; returns per start
  // (q, success, error, iterations) and the lowest converged start per target.
  py::tuple clik_multistart_raw(int64_t link, py::object target, py::object q0, int nstart, py::object mask,
                                double eps, int max_iterations, double dt, double damp, py::object stream) {
    StreamScope scope_(stream, info().device);
    Batch b(*this, target, 7);
    const int64_t total = b.B * nstart;
    py::object Q0 = b.torch.attr("as_tensor")(q0).attr("to")(py::arg("device") = b.dev,
                                                                py::arg("dtype") = b.torch.attr("float32"));
    Q0 = Q0.attr("contiguous")();
    auto vq0 = take(Q0, dlabi::kFloat, {total, b.n}, b.device);
    py::object Q = b.empty({total, b.n}), E = b.empty({total, 6});
    py::object Sx = b.empty_i32({total}), I = b.empty_i32({total}), solved = b.empty_i32({b.B});
    py::object maskT = py::none();
    std::unique_ptr<DLIn> vmask;
    if (!mask.is_none()) {
      maskT = b.input_i32(mask, b.n);
      vmask = take(maskT, dlabi::kInt, {b.n}, b.device);
    }
    auto vt = b.view(b.q), vQ = b.view(Q), vE = b.view(E);
    auto vS = take(Sx, dlabi::kInt, {total}, b.device), vI = take(I, dlabi::kInt, {total}, b.device),
         vs = take(solved, dlabi::kInt, {b.B}, b.device);
    check(sapamd_clik_multistart_batched(m_, b.B, nstart, (int32_t)clamp_index(link), f(vt), f(vq0),
                                         vmask ? static_cast<int32_t *>(vmask->ptr) : nullptr, eps, max_iterations,
                                         dt, damp, f(vQ), static_cast<int32_t *>(vS->ptr), f(vE),
                                         static_cast<int32_t *>(vI->ptr), static_cast<int32_t *>(vs->ptr),
                                         b.stream(stream)));
    return py::make_tuple(Q, Sx, E, I, solved);
  }

  // ---------------------------------------------------------------- adjoints (DLPack, FP32)
  py::tuple kin_vjp_batched(py::object q, py::object qd, py::object qdd, py::object gposes, py::object gtau,
                            py::object stream) {
    StreamScope scope_(stream, info().device);
    Batch b(*this, q);
    const bool rnea = !gtau.is_none();
    py::object qdT, qddT, gT, gP;
    std::unique_ptr<DLIn> vqd, vqdd, vgt, vgp;
    if (rnea) {
      qdT = b.input(qd);
      qddT = b.input(qdd);
      gT = b.input(gtau);
      vqd = b.view(qdT);
      vqdd = b.view(qddT);
      vgt = b.view(gT);
    }
    if (!gposes.is_none()) {
      int L = info().nlinks;
      gP = b.as_shape(gposes, {b.B, L, 7});
      vgp = b.view(gP);
    }
    py::object dq = b.empty({b.B, b.n}), dqd = py::none(), dqdd = py::none();
    std::unique_ptr<DLIn> vdqd, vdqdd;
    if (rnea) {
      dqd = b.empty({b.B, b.n});
      dqdd = b.empty({b.B, b.n});
      vdqd = b.view(dqd);
      vdqdd = b.view(dqdd);
    }
    auto vq = b.view(b.q), vdq = b.view(dq);
    check(sapamd_kin_vjp_batched(m_, b.B, f(vq), rnea ? f(vqd) : nullptr, rnea ? f(vqdd) : nullptr,
                                 vgp ? f(vgp) : nullptr, rnea ? f(vgt) : nullptr, f(vdq), rnea ? f(vdqd) : nullptr,
                                 rnea ? f(vdqdd) : nullptr, b.stream(stream)));
    return py::make_tuple(dq, dqd, dqdd);
  }
  py::tuple pose_torque_loss_batched(py::object q, py::object qd, py::object qdd, py::object tau_star,
                                     int64_t ee_link, py::object p_target, py::object stream) {
    StreamScope scope_(stream, info().device);
    Batch b(*this, q);
    const bool rnea = !tau_star.is_none();
    py::object qdT, qddT, tsT, ptT;
    std::unique_ptr<DLIn> vqd, vqdd, vts, vpt;
    if (rnea) {
      qdT = b.input(qd);
      qddT = b.input(qdd);
      tsT = b.input(tau_star);
      vqd = b.view(qdT);
      vqdd = b.view(qddT);
      vts = b.view(tsT);
    }
    if (!p_target.is_none()) {
      ptT = b.input(p_target, 3);
      vpt = b.view(ptT);
    }
    py::object loss = b.empty({b.B}), tau = py::none(), dq = b.empty({b.B, b.n}), dqd = py::none(), dqdd = py::none();
    std::unique_ptr<DLIn> vtau, vdqd, vdqdd;
    if (rnea) {
      tau = b.empty({b.B, b.n});
      dqd = b.empty({b.B, b.n});
      dqdd = b.empty({b.B, b.n});
      vtau = b.view(tau);
      vdqd = b.view(dqd);
      vdqdd = b.view(dqdd);
    }
    auto vq = b.view(b.q), vl = b.view(loss), vdq = b.view(dq);
    check(sapamd_pose_torque_loss_batched(
        m_, b.B, f(vq), rnea ? f(vqd) : nullptr, rnea ? f(vqdd) : nullptr, rnea ? f(vts) : nullptr,
        (int32_t)clamp_index(ee_link), vpt ? f(vpt) : nullptr, f(vl), rnea ? f(vtau) : nullptr, f(vdq),
        rnea ? f(vdqd) : nullptr, rnea ? f(vdqdd) : nullptr, b.stream(stream)));
    return py::make_tuple(loss, tau, dq, dqd, dqdd);
  }

  // ---------------------------------------------------------------- rollouts (DLPack, FP32)
  // (T, K) of a (B, T, nv) torque tensor and a checkpoint interval
  static int32_t ckpt_count(int64_t T, int64_t K) { return (int32_t)((T + K - 1) / K); }
  py::tuple rollout_batched(py::object q0, py::object qd0, py::object tau, double dt, int64_t K, py::object stream) {
    StreamScope scope_(stream, info().device);
    if (K < 1) throw py::value_error("checkpoint_every must be >= 1");
    Batch b(*this, q0);
    py::object qdT = b.input(qd0);
    auto shp = b.torch.attr("as_tensor")(tau).attr("shape").cast<std::vector<int64_t>>();
    if (shp.size() != 3 || shp[0] != b.B || shp[2] != b.n)
      throw std::runtime_error("tau must have shape (B, T, " + std::to_string(b.n) + ")");
    const int64_t T = shp[1];
    py::object tauT = b.as_shape(tau, shp);
    py::object Q = b.q.attr("clone")(), QD = qdT.attr("clone")();
    py::object C = b.empty({b.B, T ? ckpt_count(T, K) : 0, 2 * b.n});
    auto vQ = b.view(Q), vQD = b.view(QD), vt = b.view(tauT), vC = b.view(C);
    check(sapamd_rollout_forward(m_, b.B, (int32_t)T, (int32_t)K, (float)dt, f(vQ), f(vQD), f(vt), f(vC),
                                 b.stream(stream)));
    return py::make_tuple(Q, QD, C);
  }
  py::tuple rollout_backward_batched(py::object tau, py::object ckpt, double dt, py::object gq, py::object gqd,
                                     int64_t K, py::object stream) {
    StreamScope scope_(stream, info().device);
    if (K < 1) throw py::value_error("checkpoint_every must be >= 1");
    Batch b(*this, gq);
    py::object aq = b.q.attr("clone")(), aqd = b.input(gqd).attr("clone")();
    auto shp = b.torch.attr("as_tensor")(tau).attr("shape").cast<std::vector<int64_t>>();
    if (shp.size() != 3 || shp[0] != b.B || shp[2] != b.n)
      throw std::runtime_error("tau must have shape (B, T, " + std::to_string(b.n) + ")");
    const int64_t T = shp[1];
    py::object tauT = b.as_shape(tau, shp);
    py::object C = b.as_shape(ckpt, {b.B, T ? ckpt_count(T, K) : 0, 2 * b.n});
    py::object G = b.torch.attr("zeros")(py::cast(shp), py::arg("device") = b.dev,
                                         py::arg("dtype") = b.torch.attr("float32"));
    py::object seg = b.empty({b.B, K, 2 * b.n}), qt = b.empty({b.B, b.n}), qdt = b.empty({b.B, b.n});
    auto vt = b.view(tauT), vC = b.view(C), vG = b.view(G), vs = b.view(seg), vqt = b.view(qt), vqdt = b.view(qdt),
         vaq = b.view(aq), vaqd = b.view(aqd);
    check(sapamd_rollout_backward(m_, b.B, (int32_t)T, (int32_t)K, (float)dt, f(vt), f(vC), f(vs), f(vqt), f(vqdt),
                                  f(vaq), f(vaqd), f(vG), b.stream(stream)));
    return py::make_tuple(aq, aqd, G);
  }
  // In place on q, qd (contiguous float32 (B, nv) tensors on the model's GPU):
  // no copies, no allocation, so a caller can capture it in a CUDA/HIP graph.
  void rollout_step_(py::object q, py::object qd, py::object tau_t, double dt, py::object stream) {
    StreamScope scope_(stream, info().device);
    Batch b(*this, q);
    for (const py::object *t : {&q, &qd, &tau_t}) {
      if (!py::isinstance(*t, b.torch.attr("Tensor")) || !t->attr("is_contiguous")().cast<bool>() ||
          !t->attr("dtype").equal(b.torch.attr("float32")) || !t->attr("is_cuda").cast<bool>())
        throw std::runtime_error("rollout_step_ needs contiguous float32 GPU tensors");
    }
    auto vq = b.view(q), vqd = take(qd, dlabi::kFloat, {b.B, b.n}, b.device),
         vt = take(tau_t, dlabi::kFloat, {b.B, b.n}, b.device);
    check(sapamd_rollout_step(m_, b.B, (float)dt, f(vq), f(vqd), f(vt), b.n, b.stream(stream)));
  }

  sapamd_model *raw() const { return m_; }

  // Heterogeneous batch: several models, one launch (FK + J + RNEA + CRBA).
  static py::list hetero_batched(py::list models, py::list qs, py::object qds, py::object qdds, py::object stream) {
    StreamScope scope_(stream, models.size() ? models[0].cast<PyModel &>().info().device : 0);
    const size_t n = models.size();
    if (qs.size() != n) throw py::value_error("one q batch per model");
    std::vector<std::unique_ptr<Batch>> batches;
    std::vector<py::object> keep;
    std::vector<std::unique_ptr<DLIn>> views;
    std::vector<const sapamd_model *> ms;
    std::vector<int64_t> Bs;
    std::vector<const float *> q, qd, qdd;
    std::vector<float *> P, J, T, M;
    py::list out;
    auto fp = [&](py::object t) {
      views.push_back(take(t, dlabi::kFloat, t.attr("shape").cast<std::vector<int64_t>>(),
                           batches.back()->device));
      return static_cast<float *>(views.back()->ptr);
    };
    void *s = nullptr;
    for (size_t i = 0; i < n; ++i) {
      PyModel &pm = models[i].cast<PyModel &>();
      batches.push_back(std::make_unique<Batch>(pm, qs[i]));
      Batch &b = *batches.back();
      if (i == 0) s = b.stream(stream);
      const int L = pm.info().nlinks;
      py::object qdT = qds.is_none() ? py::object(b.torch.attr("zeros_like")(b.q)) : b.input(qds.cast<py::list>()[i]);
      py::object qddT =
          qdds.is_none() ? py::object(b.torch.attr("zeros_like")(b.q)) : b.input(qdds.cast<py::list>()[i]);
      py::object Pt = b.empty({b.B, L, 7}), Jt = b.empty({b.B, 6, b.n}), Tt = b.empty({b.B, b.n}),
                 Mt = b.empty({b.B, b.n, b.n});
      keep.insert(keep.end(), {b.q, qdT, qddT});
      ms.push_back(pm.raw());
      Bs.push_back(b.B);
      q.push_back(fp(b.q));
      qd.push_back(fp(qdT));
      qdd.push_back(fp(qddT));
      P.push_back(fp(Pt));
      J.push_back(fp(Jt));
      T.push_back(fp(Tt));
      M.push_back(fp(Mt));
      py::dict d;
      d["poses"] = Pt;
      d["J"] = Jt;
      d["tau"] = Tt;
      d["M"] = Mt;
      out.append(d);
    }
    check(sapamd_hetero_batched(ms.data(), (int32_t)n, Bs.data(), q.data(), qd.data(), qdd.data(), P.data(),
                                J.data(), T.data(), M.data(), s));
    return out;
  }

  // Heterogeneous CLIK: several models' problem batches in one persistent launch.
  static py::list clik_hetero_batched(py::list models, py::list links, py::list poses, py::object q0s,
                                      py::object masks, double eps, int max_iterations, double dt, double damp,
                                      py::object stream) {
    StreamScope scope_(stream, models.size() ? models[0].cast<PyModel &>().info().device : 0);
    const size_t n = models.size();
    if (links.size() != n || poses.size() != n) throw py::value_error("one link and one pose batch per model");
    std::vector<std::unique_ptr<Batch>> batches;
    std::vector<py::object> keep;
    std::vector<std::unique_ptr<DLIn>> views;
    std::vector<const sapamd_model *> ms;
    std::vector<int64_t> Bs;
    std::vector<int32_t> lk;
    std::vector<const float *> tg, q0;
    std::vector<const int32_t *> mk;
    std::vector<float *> Q, E;
    std::vector<int32_t *> Sx, It;
    bool any_q0 = false, any_mask = false;
    py::list out;
    std::vector<std::array<py::object, 4>> res;
    void *s = nullptr;
    for (size_t i = 0; i < n; ++i) {
      PyModel &pm = models[i].cast<PyModel &>();
      batches.push_back(std::make_unique<Batch>(pm, poses[i], 7));
      Batch &b = *batches.back();
      if (i == 0) s = b.stream(stream);
      ms.push_back(pm.raw());
      Bs.push_back(b.B);
      lk.push_back((int32_t)clamp_index(links[i].cast<int64_t>()));
      keep.push_back(b.q);
      views.push_back(b.view(b.q));
      tg.push_back(f(views.back()));
      py::object q0i = q0s.is_none() ? py::none() : py::object(q0s.cast<py::list>()[i]);
      if (!q0i.is_none()) {
        any_q0 = true;
        py::object t = b.input(q0i);
        keep.push_back(t);
        views.push_back(b.view(t));
        q0.push_back(f(views.back()));
      } else {
        q0.push_back(nullptr);
      }
      py::object mi = masks.is_none() ? py::none() : py::object(masks.cast<py::list>()[i]);
      if (!mi.is_none()) {
        any_mask = true;
        py::object t = b.input_i32(mi, b.n);
        keep.push_back(t);
        views.push_back(take(t, dlabi::kInt, {b.n}, b.device));
        mk.push_back(static_cast<const int32_t *>(views.back()->ptr));
      } else {
        mk.push_back(nullptr);
      }
      py::object Qt = b.empty({b.B, b.n}), Et = b.empty({b.B, 6}), St = b.empty_i32({b.B}), It_ = b.empty_i32({b.B});
      views.push_back(b.view(Qt));
      Q.push_back(f(views.back()));
      views.push_back(b.view(Et));
      E.push_back(f(views.back()));
      views.push_back(take(St, dlabi::kInt, {b.B}, b.device));
      Sx.push_back(static_cast<int32_t *>(views.back()->ptr));
      views.push_back(take(It_, dlabi::kInt, {b.B}, b.device));
      It.push_back(static_cast<int32_t *>(views.back()->ptr));
      res.push_back({Qt, St, Et, It_});
    }
    check(sapamd_clik_hetero_batched(ms.data(), (int32_t)n, Bs.data(), lk.data(), tg.data(),
                                     any_q0 ? q0.data() : nullptr, any_mask ? mk.data() : nullptr, eps,
                                     max_iterations, dt, damp, Q.data(), Sx.data(), E.data(), It.data(), s));
    // the int32 -> bool conversion is stream-ordered after the launch
    for (auto &r : res) out.append(py::make_tuple(r[0], r[1].attr("bool")(), r[2], r[3]));
    return out;
  }

  // Mixed-tree rollouts: several models' batches in one launch (forward, then
  // its adjoint), per model as rollout_batched / rollout_backward_batched.
  template <class BatchT> static int64_t rollout_steps(BatchT &b, py::object tau) {
    auto shp = b.torch.attr("as_tensor")(tau).attr("shape").template cast<std::vector<int64_t>>();
    if (shp.size() != 3 || shp[0] != b.B || shp[2] != b.n)
      throw std::runtime_error("tau must have shape (B, T, " + std::to_string(b.n) + ")");
    return shp[1];
  }
  static py::list rollout_hetero_batched(py::list models, py::list q0s, py::list qd0s, py::list taus, double dt,
                                         int64_t K, py::object stream) {
    StreamScope scope_(stream, models.size() ? models[0].cast<PyModel &>().info().device : 0);
    const size_t n = models.size();
    if (q0s.size() != n || qd0s.size() != n || taus.size() != n)
      throw py::value_error("one qpos, qvel and qf batch per model");
    if (K < 1) throw py::value_error("checkpoint_every must be >= 1");
    std::vector<std::unique_ptr<Batch>> batches;
    std::vector<std::unique_ptr<DLIn>> views;
    std::vector<py::object> keep;
    std::vector<const sapamd_model *> ms;
    std::vector<int64_t> Bs;
    std::vector<float *> Q, QD, C;
    std::vector<const float *> tau;
    py::list out;
    int64_t T = -1;
    void *s = nullptr;
    for (size_t i = 0; i < n; ++i) {
      PyModel &pm = models[i].cast<PyModel &>();
      batches.push_back(std::make_unique<Batch>(pm, q0s[i]));
      Batch &b = *batches.back();
      if (i == 0) s = b.stream(stream);
      const int64_t Ti = rollout_steps(b, taus[i]);
      if (T >= 0 && Ti != T) throw std::runtime_error("every model's tau needs the same number of steps");
      T = Ti;
      py::object tauT = b.as_shape(taus[i], {b.B, T, b.n});
      py::object Qt = b.q.attr("clone")(), QDt = b.input(qd0s[i]).attr("clone")();
      py::object Ct = b.empty({b.B, T ? ckpt_count(T, K) : 0, 2 * b.n});
      keep.push_back(tauT);
      for (py::object *t : {&Qt, &QDt, &tauT, &Ct}) views.push_back(b.view(*t));
      Q.push_back(f(views[views.size() - 4]));
      QD.push_back(f(views[views.size() - 3]));
      tau.push_back(f(views[views.size() - 2]));
      C.push_back(f(views[views.size() - 1]));
      ms.push_back(pm.raw());
      Bs.push_back(b.B);
      out.append(py::make_tuple(Qt, QDt, Ct));
    }
    if (n)
      check(sapamd_rollout_forward_hetero(ms.data(), (int32_t)n, Bs.data(), (int32_t)T, (int32_t)K, (float)dt,
                                          Q.data(), QD.data(), tau.data(), C.data(), s));
    return out;
  }
  static py::list rollout_backward_hetero_batched(py::list models, py::list taus, py::list ckpts, double dt,
                                                  py::list gqs, py::list gqds, int64_t K, py::object stream) {
    StreamScope scope_(stream, models.size() ? models[0].cast<PyModel &>().info().device : 0);
    const size_t n = models.size();
    if (taus.size() != n || ckpts.size() != n || gqs.size() != n || gqds.size() != n)
      throw py::value_error("one qf, checkpoint and gradient batch per model");
    if (K < 1) throw py::value_error("checkpoint_every must be >= 1");
    std::vector<std::unique_ptr<Batch>> batches;
    std::vector<std::unique_ptr<DLIn>> views;
    std::vector<py::object> keep;
    std::vector<const sapamd_model *> ms;
    std::vector<int64_t> Bs;
    std::vector<const float *> tau, C;
    std::vector<float *> SG, AQ, AQD, GT;
    py::list out;
    int64_t T = -1;
    void *s = nullptr;
    for (size_t i = 0; i < n; ++i) {
      PyModel &pm = models[i].cast<PyModel &>();
      batches.push_back(std::make_unique<Batch>(pm, gqs[i]));
      Batch &b = *batches.back();
      if (i == 0) s = b.stream(stream);
      const int64_t Ti = rollout_steps(b, taus[i]);
      if (T >= 0 && Ti != T) throw std::runtime_error("every model's tau needs the same number of steps");
      T = Ti;
      py::object tauT = b.as_shape(taus[i], {b.B, T, b.n});
      py::object Ct = b.as_shape(ckpts[i], {b.B, T ? ckpt_count(T, K) : 0, 2 * b.n});
      py::object aq = b.q.attr("clone")(), aqd = b.input(gqds[i]).attr("clone")();
      py::object Gt = b.torch.attr("zeros")(py::make_tuple(b.B, T, b.n), py::arg("device") = b.dev,
                                            py::arg("dtype") = b.torch.attr("float32"));
      py::object seg = b.empty({b.B, K, 2 * b.n});
      keep.insert(keep.end(), {tauT, Ct, seg});
      for (py::object *t : {&tauT, &Ct, &seg, &aq, &aqd, &Gt}) views.push_back(b.view(*t));
      const size_t v = views.size() - 6;
      tau.push_back(f(views[v]));
      C.push_back(f(views[v + 1]));
      SG.push_back(f(views[v + 2]));
      AQ.push_back(f(views[v + 3]));
      AQD.push_back(f(views[v + 4]));
      GT.push_back(f(views[v + 5]));
      ms.push_back(pm.raw());
      Bs.push_back(b.B);
      out.append(py::make_tuple(aq, aqd, Gt));
    }
    if (n)
      check(sapamd_rollout_backward_hetero(ms.data(), (int32_t)n, Bs.data(), (int32_t)T, (int32_t)K, (float)dt,
                                           tau.data(), C.data(), SG.data(), AQ.data(), AQD.data(), GT.data(), s));
    return out;
  }

 private:
  sapamd_model *m_ = nullptr;
  std::mt19937_64 rng_{std::random_device{}()};

  static int64_t clamp_index(int64_t i) { return i < 0 || i > INT32_MAX ? INT32_MAX : i; }
  static void same_len(const DArr &a, const DArr &b) {
    if (a.size() != b.size()) throw std::runtime_error("qpos, qvel, qacc/qf must have the same length");
  }
  static float *f(const std::unique_ptr<DLIn> &v) { return static_cast<float *>(v->ptr); }

  // Normalizes batched inputs: any tensor / array-like becomes a contiguous
  // float32 tensor on the model's GPU, shape (B, ncols).
  // With a stream given, a batched call stages its inputs (device and dtype
  // conversion) and allocates its outputs with that stream as torch's current
  // stream, so those copies precede the kernel on the same stream and the
  // caching allocator ties the temporaries to it.
  struct StreamScope {
    py::object ctx;
    StreamScope(const py::object &stream, int device) {
      if (stream.is_none()) return;
      py::module_ torch = py::module_::import("torch");
      py::object ext = torch.attr("cuda").attr("ExternalStream")(
          stream.cast<intptr_t>(), py::arg("device") = torch.attr("device")("cuda", device));
      ctx = torch.attr("cuda").attr("stream")(ext);
      ctx.attr("__enter__")();
    }
    ~StreamScope() {
      if (!ctx) return;
      try {
        ctx.attr("__exit__")(py::none(), py::none(), py::none());
      } catch (...) {
      }
    }
    StreamScope(const StreamScope &) = delete;
    StreamScope &operator=(const StreamScope &) = delete;
  };

  struct Batch {
    PyModel &self;
    py::module_ torch;
    py::object dev, q;
    int device;
    int64_t B = 0, n = 0;
    Batch(PyModel &s, py::object q_in, int64_t ncols = -1) : self(s), torch(py::module_::import("torch")) {
      sapamd_info i = s.info();
      device = i.device;
      n = i.nv;
      dev = torch.attr("device")("cuda", device);
      q = input(q_in, ncols < 0 ? n : ncols);
      B = q.attr("shape").cast<py::tuple>()[0].cast<int64_t>();
    }
    py::object input(py::object x, int64_t ncols = -1) {
      if (ncols < 0) ncols = n;
      if (x.is_none()) throw py::value_error("missing batched input");
      if (!py::isinstance(x, torch.attr("Tensor"))) x = torch.attr("as_tensor")(x);
      x = x.attr("detach")().attr("to")(py::arg("device") = dev, py::arg("dtype") = torch.attr("float32"));
      if (x.attr("dim")().cast<int>() == 1) x = x.attr("unsqueeze")(0);
      x = x.attr("contiguous")();
      auto shp = x.attr("shape").cast<py::tuple>();
      if (shp.size() != 2 || shp[1].cast<int64_t>() != ncols)
        throw std::runtime_error("batched input must have shape (B, " + std::to_string(ncols) + ")");
      if (B && shp[0].cast<int64_t>() != B) throw std::runtime_error("batched inputs disagree on B");
      return x;
    }
    py::object as_shape(py::object x, std::vector<int64_t> shape) {
      if (!py::isinstance(x, torch.attr("Tensor"))) x = torch.attr("as_tensor")(x);
      x = x.attr("detach")().attr("to")(py::arg("device") = dev, py::arg("dtype") = torch.attr("float32"));
      x = x.attr("contiguous")();
      if (x.attr("shape").cast<std::vector<int64_t>>() != shape) throw std::runtime_error("unexpected gradient shape");
      return x;
    }
    py::object input_i32(py::object x, int64_t len) {
      if (!py::isinstance(x, torch.attr("Tensor"))) x = torch.attr("as_tensor")(x);
      x = x.attr("detach")().attr("to")(py::arg("device") = dev, py::arg("dtype") = torch.attr("int32")).attr("contiguous")();
      if (x.attr("numel")().cast<int64_t>() != len) throw std::runtime_error("active_qmask must have length nv");
      return x.attr("reshape")(len);
    }
    py::object empty(std::vector<int64_t> shape) {
      return torch.attr("empty")(py::cast(shape), py::arg("device") = dev, py::arg("dtype") = torch.attr("float32"));
    }
    py::object empty_i32(std::vector<int64_t> shape) {
      return torch.attr("empty")(py::cast(shape), py::arg("device") = dev, py::arg("dtype") = torch.attr("int32"));
    }
    std::unique_ptr<DLIn> view(const py::object &t) {
      auto shp = t.attr("shape").cast<std::vector<int64_t>>();
      return take(t, dlabi::kFloat, shp, device);
    }
    void *stream(const py::object &s) {
      if (!s.is_none()) return reinterpret_cast<void *>(s.cast<intptr_t>());
      py::object cs = torch.attr("cuda").attr("current_stream")(dev);
      return reinterpret_cast<void *>(cs.attr("cuda_stream").cast<intptr_t>());
    }
  };
};

}  // namespace

PYBIND11_MODULE(_pinocchio, m) {
  m.doc() = "MI355X-native PinocchioModel (HIP kernels behind a C ABI)";
  m.def("version", [] { return std::string(sapamd_version()); });
  m.def("_abi_handle", [](PyModel &pm) { return reinterpret_cast<intptr_t>(pm.raw()); },
        "raw sapamd_model* of a model, for ctypes callers of the C ABI");
  m.def("hetero_batched", &PyModel::hetero_batched, py::arg("models"), py::arg("qpos"), py::arg("qvel") = py::none(),
        py::arg("qacc") = py::none(), py::arg("stream") = py::none(),
        "Heterogeneous batch of several PinocchioModels (trees of <= 32 joints) in ONE kernel launch: per model "
        "a dict of link poses (B, L, 7), WORLD Jacobian (B, 6, nv), RNEA torques (B, nv) and mass matrix (B, nv, nv).");

  m.def("rollout_hetero_batched", &PyModel::rollout_hetero_batched, py::arg("models"), py::arg("qpos"),
        py::arg("qvel"), py::arg("qf"), py::arg("dt"), py::arg("checkpoint_every") = 16, py::arg("stream") = py::none(),
        "Rollouts of several PinocchioModels (trees of <= 32 joints) in ONE launch: per model (q_T, qd_T, "
        "checkpoints), as rollout_batched.");
  m.def("rollout_backward_hetero_batched", &PyModel::rollout_backward_hetero_batched, py::arg("models"),
        py::arg("qf"), py::arg("checkpoints"), py::arg("dt"), py::arg("grad_qpos"), py::arg("grad_qvel"),
        py::arg("checkpoint_every") = 16, py::arg("stream") = py::none(),
        "Adjoints of rollout_hetero_batched in ONE launch: per model (dL/dq0, dL/dqd0, dL/dqf).");

  m.def("clik_hetero_batched", &PyModel::clik_hetero_batched, py::arg("models"), py::arg("link_indices"),
        py::arg("poses"), py::arg("initial_qpos") = py::none(), py::arg("active_qmask") = py::none(),
        py::arg("eps") = 1e-4, py::arg("max_iterations") = 1000, py::arg("dt") = 0.1, py::arg("damp") = 1e-6,
        py::arg("stream") = py::none(),
        "CLIK for several PinocchioModels in ONE persistent launch: per model (qpos (B, nv), success (B,), "
        "error (B, 6), iterations (B,)).");

  py::class_<Pose>(m, "Pose")
      .def(py::init([](std::vector<float> p, std::vector<float> q) {
             if (p.size() != 3 || q.size() != 4) throw py::value_error("Pose(p: 3, q: 4)");
             Pose r;
             for (int i = 0; i < 3; ++i) r.p[i] = p[i];
             for (int i = 0; i < 4; ++i) r.q[i] = q[i];
             return r;
           }),
           py::arg("p") = std::vector<float>{0, 0, 0}, py::arg("q") = std::vector<float>{1, 0, 0, 0})
      .def(py::init([](const DArr &T) {
             if (T.ndim() != 2 || T.shape(0) != 4 || T.shape(1) != 4) throw py::value_error("expected a 4x4 matrix");
             const double *t = T.data();
             double R[9] = {t[0], t[1], t[2], t[4], t[5], t[6], t[8], t[9], t[10]};
             Pose r;
             r.p = {float(t[3]), float(t[7]), float(t[11])};
             r.q = Pose::quat_of(R);
             return r;
           }),
           py::arg("matrix"))
      .def_property(
          "p", [](const Pose &s) { return py::array_t<float>(3, s.p.data()); },
          [](Pose &s, std::vector<float> v) { for (int i = 0; i < 3; ++i) s.p[i] = v.at(i); })
      .def_property(
          "q", [](const Pose &s) { return py::array_t<float>(4, s.q.data()); },
          [](Pose &s, std::vector<float> v) { for (int i = 0; i < 4; ++i) s.q[i] = v.at(i); })
      .def("get_p", [](const Pose &s) { return py::array_t<float>(3, s.p.data()); })
      .def("get_q", [](const Pose &s) { return py::array_t<float>(4, s.q.data()); })
      .def("set_p", [](Pose &s, std::vector<float> v) { for (int i = 0; i < 3; ++i) s.p[i] = v.at(i); })
      .def("set_q", [](Pose &s, std::vector<float> v) { for (int i = 0; i < 4; ++i) s.q[i] = v.at(i); })
      .def_property(
          "rpy", [](const Pose &s) { auto r = s.rpy(); return py::array_t<float>(3, r.data()); },
          [](Pose &s, std::vector<float> v) { s.set_rpy({v.at(0), v.at(1), v.at(2)}); })
      .def("get_rpy", [](const Pose &s) { auto r = s.rpy(); return py::array_t<float>(3, r.data()); })
      .def("set_rpy", [](Pose &s, std::vector<float> v) { s.set_rpy({v.at(0), v.at(1), v.at(2)}); })
      .def("inv", &Pose::inv)
      .def("__mul__", &Pose::mul)
      .def("to_transformation_matrix",
           [](const Pose &s) {
             auto R = s.matrix();
             py::array_t<float> T({4, 4});
             float *t = T.mutable_data();
             for (int i = 0; i < 16; ++i) t[i] = 0;
             for (int i = 0; i < 3; ++i) {
               for (int j = 0; j < 3; ++j) t[4 * i + j] = float(R[3 * i + j]);
               t[4 * i + 3] = s.p[i];
             }
             t[15] = 1;
             return T;
           })
      .def(py::pickle([](const Pose &s) { return py::make_tuple(s.p, s.q); },
                      [](py::tuple t) {
                        Pose r;
                        r.p = t[0].cast<std::array<float, 3>>();
                        r.q = t[1].cast<std::array<float, 4>>();
                        return r;
                      }))
      .def("__repr__", [](const Pose &s) {
        std::ostringstream os;
        os << "Pose([" << s.p[0] << ", " << s.p[1] << ", " << s.p[2] << "], [" << s.q[0] << ", " << s.q[1] << ", "
           << s.q[2] << ", " << s.q[3] << "])";
        return os.str();
      });

  py::class_<PyModel>(m, "PinocchioModel")
      .def(py::init<const std::string &, const std::vector<double> &>(), py::arg("urdf"), py::arg("gravity"))
      .def("set_link_order", &PyModel::set_link_order)
      .def("set_joint_order", &PyModel::set_joint_order)
      .def("compute_forward_kinematics", &PyModel::compute_forward_kinematics,
           "Compute and cache forward kinematics. After computation, use get_link_pose to retrieve the computed "
           "pose for a specific link.",
           py::arg("qpos"))
      .def("get_link_pose", &PyModel::get_link_pose,
           "Given link index, get link pose (in articulation base frame) from forward kinematics. Must be called "
           "after compute_forward_kinematics.",
           py::arg("link_index"))
      .def("compute_inverse_kinematics", &PyModel::compute_inverse_kinematics,
           "Compute inverse kinematics with the CLIK algorithm (damped least squares on the SE(3) log error).\n"
           "Returns (qpos, success, error6).",
           py::arg("link_index"), py::arg("pose"), py::arg("initial_qpos") = DArr(), py::arg("active_qmask") = IArr(),
           py::arg("eps") = 1e-4, py::arg("max_iterations") = 1000, py::arg("dt") = 0.1, py::arg("damp") = 1e-6)
      .def("compute_forward_dynamics", &PyModel::compute_forward_dynamics, py::arg("qpos"), py::arg("qvel"),
           py::arg("qf"))
      .def("compute_inverse_dynamics", &PyModel::compute_inverse_dynamics, py::arg("qpos"), py::arg("qvel"),
           py::arg("qacc"))
      .def("compute_generalized_mass_matrix", &PyModel::compute_generalized_mass_matrix, py::arg("qpos"))
      .def("compute_coriolis_matrix", &PyModel::compute_coriolis_matrix, py::arg("qpos"), py::arg("qvel"))
      .def("compute_full_jacobian", &PyModel::compute_full_jacobian, "Compute and cache Jacobian for all links",
           py::arg("qpos"))
      .def("get_link_jacobian", &PyModel::get_link_jacobian,
           "Given link index, get the Jacobian. Must be called after compute_full_jacobian.\n"
           "local=False: world (spatial) frame; local=True: link (body) frame.",
           py::arg("link_index"), py::arg("local") = false)
      .def("compute_single_link_local_jacobian", &PyModel::compute_single_link_local_jacobian,
           "Compute the link(body) Jacobian for a single link.", py::arg("qpos"), py::arg("link_index"))
      .def("get_random_qpos", &PyModel::get_random_qpos)
      .def("get_random_qpos_batched", &PyModel::random_qpos_batched, py::arg("batch_size"),
           py::arg("seed") = py::none(), py::arg("stream") = py::none(),
           "(B, nv) FP32 configurations uniform within the joint limits (continuous joints in [-pi, pi)), "
           "Philox4x32-10 at counter (b, u) under the seed: the same seed gives the same batch.")
      .def("joint_limits", &PyModel::joint_limits)
      .def_property_readonly("nv", [](const PyModel &s) { return s.info().nv; })
      .def_property_readonly("nq", [](const PyModel &s) { return s.info().nq; })
      .def_property_readonly("nlinks", [](const PyModel &s) { return s.info().nlinks; })
      .def_property_readonly("is_chain", [](const PyModel &s) { return bool(s.info().chain); })
      .def_property_readonly("max_depth", [](const PyModel &s) { return s.info().max_depth; })
      // batched FP32 methods over (B, dof) tensors (DLPack, kDLROCM)
      .def("compute_forward_kinematics_batched", &PyModel::fk_batched, py::arg("qpos"), py::arg("stream") = py::none(),
           "Link poses (B, L, 7) = p xyz | q wxyz for a (B, nv) batch.")
      .def("compute_full_jacobian_batched", &PyModel::jac_batched, py::arg("qpos"), py::arg("stream") = py::none(),
           "WORLD-frame Jacobian (B, 6, nv).")
      .def("compute_link_jacobians_batched", &PyModel::link_jacobians_batched, py::arg("qpos"),
           py::arg("stream") = py::none(), "Per-link Jacobians (world, local), each (B, L, 6, nv).")
      .def("compute_inverse_dynamics_batched", &PyModel::rnea_batched, py::arg("qpos"), py::arg("qvel"),
           py::arg("qacc"), py::arg("stream") = py::none())
      .def("compute_fk_jacobian_rnea_batched",
           [](PyModel &s, py::object q, py::object qd, py::object qdd, py::object stream) {
             return s.fk_jac_rnea_batched(q, qd, qdd, true, true, true, stream);
           },
           py::arg("qpos"), py::arg("qvel"), py::arg("qacc"), py::arg("stream") = py::none(),
           "Fused link poses (B, L, 7), Jacobian (B, 6, nv) and RNEA torques (B, nv).")
      .def("compute_generalized_mass_matrix_batched", &PyModel::crba_batched, py::arg("qpos"),
           py::arg("stream") = py::none())
      .def("compute_coriolis_matrix_batched", &PyModel::coriolis_batched, py::arg("qpos"), py::arg("qvel"),
           py::arg("stream") = py::none())
      .def("compute_forward_dynamics_batched", &PyModel::aba_batched, py::arg("qpos"), py::arg("qvel"), py::arg("qf"),
           py::arg("stream") = py::none())
      .def("kinematics_vjp_batched", &PyModel::kin_vjp_batched, py::arg("qpos"), py::arg("qvel") = py::none(),
           py::arg("qacc") = py::none(), py::arg("grad_poses") = py::none(), py::arg("grad_tau") = py::none(),
           py::arg("stream") = py::none(),
           "Reverse mode of link poses and RNEA: (dq, dqd, dqdd) for upstream gradients on poses (B, L, 7) "
           "and/or torques (B, nv).")
      .def("rollout_batched", &PyModel::rollout_batched, py::arg("qpos"), py::arg("qvel"), py::arg("qf"),
           py::arg("dt"), py::arg("checkpoint_every") = 16, py::arg("stream") = py::none(),
           "Semi-implicit Euler rollouts with ABA: qf (B, T, nv) -> final (q, qd) and checkpoints (B, ceil(T/K), 2nv)")
      .def("rollout_backward_batched", &PyModel::rollout_backward_batched, py::arg("qf"), py::arg("checkpoints"),
           py::arg("dt"), py::arg("grad_qpos"), py::arg("grad_qvel"), py::arg("checkpoint_every") = 16,
           py::arg("stream") = py::none(), "Adjoint of rollout_batched: (dL/dq0, dL/dqd0, dL/dqf)")
      .def("rollout_step_", &PyModel::rollout_step_, py::arg("qpos"), py::arg("qvel"), py::arg("qf"), py::arg("dt"),
           py::arg("stream") = py::none(), "One in-place rollout step (graph-capturable)")
      .def("compute_fk_jacobian_rnea_into", &PyModel::fk_jac_rnea_into, py::arg("qpos"), py::arg("qvel") = py::none(),
           py::arg("qacc") = py::none(), py::arg("poses") = py::none(), py::arg("jacobian") = py::none(),
           py::arg("qf") = py::none(), py::arg("stream") = py::none(),
           "Fused FK / WORLD Jacobian / RNEA into caller buffers, without copies: joint arrays (B, nv) with any "
           "row stride (e.g. [:, :nv] of a padded [n_art, max_dof] state), poses (B, >= L, >= 7) with the last "
           "dimension contiguous (e.g. a [n_art, max_links, 13] body-state buffer: p | q wxyz in the first 7).")
      .def("compute_link_state_into", &PyModel::link_state_into, py::arg("qpos"), py::arg("qvel") = py::none(),
           py::arg("body") = py::none(), py::arg("stream") = py::none(),
           "Link poses and velocities into a GPU simulator's body-state buffer (B, >= L, >= 13) rows "
           "p | q wxyz | v | w, world frame: row 0 (the root link) is read as the articulation's world pose and "
           "velocity, then every link's row is written (v = centre-of-mass velocity); without qvel only the "
           "poses. Joint arrays (B, nv) with any row stride.")
      .def("compute_passive_force_batched", &PyModel::passive_force_batched, py::arg("qpos"),
           py::arg("qvel") = py::none(), py::arg("gravity") = true, py::arg("coriolis_and_centrifugal") = true,
           py::arg("out") = py::none(), py::arg("stream") = py::none(),
           "Passive joint forces g(q) + C(q, qd) qd (either term optional), as "
           "PhysxArticulation.compute_passive_force: RNEA(q, qd, 0). Inputs may be padded (row-strided).")
      .def("_clik_multistart_raw", &PyModel::clik_multistart_raw, py::arg("link_index"), py::arg("poses"),
           py::arg("initial_qpos"), py::arg("n_start"), py::arg("active_qmask") = py::none(), py::arg("eps") = 1e-4,
           py::arg("max_iterations") = 1000, py::arg("dt") = 0.1, py::arg("damp") = 1e-6,
           py::arg("stream") = py::none())
      .def("pose_torque_loss_batched", &PyModel::pose_torque_loss_batched, py::arg("qpos"),
           py::arg("qvel") = py::none(), py::arg("qacc") = py::none(), py::arg("tau_target") = py::none(),
           py::arg("ee_link") = 0, py::arg("p_target") = py::none(), py::arg("stream") = py::none(),
           "loss = |tau(q, qd, qdd) - tau*|^2 + |p_ee(q) - t|^2 per environment with its gradients, one pass: "
           "(loss (B,), tau (B, nv), dq, dqd, dqdd).")
      .def("compute_inverse_kinematics_batched", &PyModel::clik_batched, py::arg("link_index"), py::arg("poses"),
           py::arg("initial_qpos") = py::none(), py::arg("active_qmask") = py::none(), py::arg("eps") = 1e-4,
           py::arg("max_iterations") = 1000, py::arg("dt") = 0.1, py::arg("damp") = 1e-6,
           py::arg("stream") = py::none(),
           "Batched CLIK: poses (B, 7) -> (qpos (B, nv), success (B,), error (B, 6), iterations (B,)).");
}
