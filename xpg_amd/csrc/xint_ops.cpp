// The integration kernels as PyTorch operators: TORCH_LIBRARY(xint) over the C ABI of libxpg_integrate.so
// (include/xpg_integrate.h).  Thin: each op checks its tensors, fills the C structs with their device pointers, runs
// the entry point on the current HIP stream and returns tensors from torch's caching allocator.  No host
// synchronisation, no global state; errors raise through TORCH_CHECK (RuntimeError in Python).
//
//   xint::extract              shoeboxes of a box table from a device frame store   (xpg_extract)
//   xint::pass1                pass1 of make_kernels, xpg/integrate_kernels.py:287-365 (xpg_pass1), optionally
//                              splatting the strong rows into fixed-point profile grids
//   xint::accumulate_profiles  accumulate_profiles, :367-380 (xpg_accumulate_profiles): float64 grids
//   xint::normalise_profiles   the profile normalisation of xpg/integrate.py:546-556 (xpg_normalise_profiles)
//   xint::pass2                pass2, xpg/integrate_kernels.py:382-483 (xpg_pass2)
//
// Row tables are lists in the order of ROW_KEYS (xpg_amd/ops.py), absent entries None; the detector is a (6, K, 3)
// float64 tensor of the affine panel maps (P0, Bx, By, U1, U2, UN) with optional segment map, correction tables and
// background template, plus its eight scalars; the call scalars are axis, s0, frames_phi0, osc, sigma_D, sigma_M and
// gain, with the 21 mixture weights or none.
#include <ATen/ATen.h>
#include <ATen/hip/HIPContext.h>
#include <torch/library.h>

#include "xpg_integrate.h"

namespace {

using at::Tensor;
using OptT = c10::optional<Tensor>;

constexpr double FX_SCALE = 17592186044416.0;   // 2^44: the fixed-point profile grids of the library

void *stream() { return (void *)at::hip::getCurrentHIPStream().stream(); }

void check_rc(int rc, const char *what) { TORCH_CHECK(rc == 0, "xint::", what, ": ", xpg_last_error()); }

template <typename T>
const T *dptr(const Tensor &t, at::ScalarType st, const char *name, int64_t numel = -1) {
  TORCH_CHECK(t.is_cuda(), name, ": a device tensor is expected");
  TORCH_CHECK(t.scalar_type() == st, name, ": dtype ", st, " expected, got ", t.scalar_type());
  TORCH_CHECK(t.is_contiguous(), name, ": contiguous tensor expected");
  TORCH_CHECK(numel < 0 || t.numel() == numel, name, ": ", numel, " elements expected, got ", t.numel());
  return static_cast<const T *>(t.data_ptr());
}

template <typename T>
const T *optp(const c10::List<OptT> &l, size_t i, at::ScalarType st, const char *name, int64_t numel = -1) {
  if (i >= l.size()) return nullptr;
  const OptT o = l.get(i);
  return o.has_value() ? dptr<T>(*o, st, name, numel) : nullptr;
}

// the row table (ROW_KEYS order): S, phi, ix0, iy0, iz0, fb, pxs, region, has_p, pS, pphi, nbx, nby, nbz, nbsx, nbsz
struct Rows {
  xpg_refl r{};
};

Rows make_refl(const c10::List<OptT> &rows, int64_t n, int64_t Dz, int64_t Dy, int64_t Dx) {
  TORCH_CHECK(rows.size() == 16, "rows: 16 entries (S, phi, ix0, iy0, iz0, fb, pxs, region, has_p, pS, pphi, nbx, nby, "
              "nbz, nbsx, nbsz)");
  Rows R;
  xpg_refl &r = R.r;
  r.n = (int32_t)n;
  r.Dz = (int32_t)Dz;
  r.Dy = (int32_t)Dy;
  r.Dx = (int32_t)Dx;
  r.S = optp<double>(rows, 0, at::kDouble, "S", 3 * n);
  r.phi = optp<double>(rows, 1, at::kDouble, "phi", n);
  r.ix0 = optp<int32_t>(rows, 2, at::kInt, "ix0", n);
  r.iy0 = optp<int32_t>(rows, 3, at::kInt, "iy0", n);
  r.iz0 = optp<int32_t>(rows, 4, at::kInt, "iz0", n);
  r.fb = optp<int32_t>(rows, 5, at::kInt, "fb", n);
  r.pxs = optp<int32_t>(rows, 6, at::kInt, "pxs", n);
  r.region = optp<int32_t>(rows, 7, at::kInt, "region", n);
  TORCH_CHECK(r.S && r.phi && r.ix0 && r.iy0 && r.iz0 && r.fb && r.region, "rows: S, phi, ix0, iy0, iz0, fb and region "
              "are required");
  r.has_p = optp<float>(rows, 8, at::kFloat, "has_p", n);
  r.pS = optp<double>(rows, 9, at::kDouble, "pS", 3 * n);
  r.pphi = optp<double>(rows, 10, at::kDouble, "pphi", n);
  TORCH_CHECK(!r.has_p || (r.pS && r.pphi), "rows: has_p needs pS and pphi");
  const OptT nbx = rows.get(11);
  if (nbx.has_value() && nbx->dim() == 2 && nbx->size(1) > 0) {
    const int64_t K = nbx->size(1);
    r.K = (int32_t)K;
    r.nbx = optp<float>(rows, 11, at::kFloat, "nbx", n * K);
    r.nby = optp<float>(rows, 12, at::kFloat, "nby", n * K);
    r.nbz = optp<float>(rows, 13, at::kFloat, "nbz", n * K);
    r.nbsx = optp<float>(rows, 14, at::kFloat, "nbsx", n * K);
    r.nbsz = optp<float>(rows, 15, at::kFloat, "nbsz", n * K);
    TORCH_CHECK(r.nby && r.nbz && r.nbsx && r.nbsz, "rows: the five neighbour columns go together");
  }
  return R;
}

// the detector: affine (6, K, 3) float64, then segmap (int16), gx, gy (float32, one shape), bkg (float32)
xpg_detector make_det(const c10::List<OptT> &det, at::ArrayRef<double> sc) {
  TORCH_CHECK(det.size() == 5, "det: 5 entries (affine, segmap, gx, gy, bkg)");
  TORCH_CHECK(sc.size() == 8, "det_scalars: orgx, orgy, qx, qy, wl, tsens, mu, plx");
  xpg_detector d{};
  const OptT aff = det.get(0);
  TORCH_CHECK(aff.has_value() && aff->dim() == 3 && aff->size(0) == 6 && aff->size(2) == 3,
              "det[0]: (6, K, 3) float64 panel maps P0, Bx, By, U1, U2, UN");
  const int64_t K = aff->size(1);
  const double *a = dptr<double>(*aff, at::kDouble, "det[0]");
  d.n_seg = (int32_t)K;
  d.P0 = a;
  d.Bx = a + 3 * K;
  d.By = a + 6 * K;
  d.U1 = a + 9 * K;
  d.U2 = a + 12 * K;
  d.UN = a + 15 * K;
  d.orgx = sc[0];
  d.orgy = sc[1];
  d.qx = sc[2];
  d.qy = sc[3];
  d.wl = sc[4];
  d.tsens = sc[5];
  d.mu = sc[6];
  d.plx = sc[7];
  const OptT seg = det.get(1), gx = det.get(2), gy = det.get(3), bkg = det.get(4);
  if (seg.has_value()) {
    TORCH_CHECK(seg->dim() == 2, "segmap: (ny, nx)");
    d.segmap = dptr<int16_t>(*seg, at::kShort, "segmap");
    d.seg_ny = (int32_t)seg->size(0);
    d.seg_nx = (int32_t)seg->size(1);
  }
  TORCH_CHECK(gx.has_value() == gy.has_value(), "gx and gy go together");
  if (gx.has_value()) {
    TORCH_CHECK(gx->dim() == 2 && gx->sizes() == gy->sizes(), "gx, gy: (ny, nx), one shape");
    d.gx = dptr<float>(*gx, at::kFloat, "gx");
    d.gy = dptr<float>(*gy, at::kFloat, "gy");
    d.geo_ny = (int32_t)gx->size(0);
    d.geo_nx = (int32_t)gx->size(1);
  }
  if (bkg.has_value()) {
    TORCH_CHECK(bkg->dim() == 2, "bkg: (ny, nx)");
    d.bkg = dptr<float>(*bkg, at::kFloat, "bkg");
    d.bkg_ny = (int32_t)bkg->size(0);
    d.bkg_nx = (int32_t)bkg->size(1);
  }
  return d;
}

xpg_call make_call(at::ArrayRef<double> c, c10::optional<at::ArrayRef<double>> rcw) {
  TORCH_CHECK(c.size() == 11, "call: axis (3), s0 (3), frames_phi0, osc, sigma_D, sigma_M, gain");
  xpg_call k{};
  for (int i = 0; i < 3; ++i) {
    k.axis[i] = c[i];
    k.s0[i] = c[3 + i];
  }
  k.frames_phi0 = c[6];
  k.osc = c[7];
  k.sd = c[8];
  k.sm = c[9];
  k.gain = c[10];
  if (rcw.has_value()) {
    TORCH_CHECK((int)rcw->size() == XPG_RC_K, "rcw: ", XPG_RC_K, " mixture weights");
    for (int i = 0; i < XPG_RC_K; ++i) k.rcw[i] = (float)(*rcw)[i];
    k.use_rcw = 1;
  }
  return k;
}

void check_sb(const Tensor &sb) {
  TORCH_CHECK(sb.dim() == 4, "sb: (n, Dz, Dy, Dx)");
  dptr<float>(sb, at::kFloat, "sb");
}

Tensor extract(const Tensor &frames, int64_t frame0, int64_t min_valid, int64_t flag_value, int64_t overload,
               bool use_flag_value, const c10::List<OptT> &rows, at::IntArrayRef box) {
  TORCH_CHECK(frames.dim() == 3 && frames.is_cuda() && frames.is_contiguous(), "frames: contiguous (n, ny, nx) on the "
              "device");
  TORCH_CHECK(frames.scalar_type() == at::kInt || frames.scalar_type() == at::kFloat, "frames: int32 (uint32 raw "
              "counts) or float32 (value-coded)");
  TORCH_CHECK(box.size() == 3, "box: Dz, Dy, Dx");
  const OptT S = rows.get(0);
  TORCH_CHECK(S.has_value(), "rows[0] (S) is required");
  const int64_t n = S->size(0);
  Rows R = make_refl(rows, n, box[0], box[1], box[2]);
  xpg_frames F{};
  F.data = frames.data_ptr();
  F.frame_stride = frames.size(1) * frames.size(2);
  F.ny = (int32_t)frames.size(1);
  F.nx = (int32_t)frames.size(2);
  F.frame0 = (int32_t)frame0;
  F.n_frames = (int32_t)frames.size(0);
  F.min_valid = (uint32_t)min_valid;
  F.flag_value = (uint32_t)flag_value;
  F.overload = (uint32_t)overload;
  F.use_flag_value = use_flag_value ? 1 : 0;
  F.dtype = frames.scalar_type() == at::kFloat ? 1 : 0;
  Tensor sb = at::empty({n, box[0], box[1], box[2]}, frames.options().dtype(at::kFloat));
  check_rc(xpg_extract(&F, &R.r, sb.data_ptr<float>(), stream()), "extract");
  return sb;
}

std::vector<Tensor> pass1(const Tensor &sb, const c10::List<OptT> &rows, const c10::List<OptT> &det,
                          at::ArrayRef<double> det_scalars, at::ArrayRef<double> call,
                          c10::optional<at::ArrayRef<double>> rcw, const OptT &acc_num, const OptT &acc_den) {
  check_sb(sb);
  const int64_t n = sb.size(0), Dz = sb.size(1), Dy = sb.size(2), Dx = sb.size(3);
  Rows R = make_refl(rows, n, Dz, Dy, Dx);
  const xpg_detector d = make_det(det, det_scalars);
  const xpg_call k = make_call(call, rcw);
  const auto f64 = sb.options().dtype(at::kDouble), i32 = sb.options().dtype(at::kInt);
  std::vector<Tensor> out;
  for (int i = 0; i < 14; ++i) out.push_back(at::empty({n}, f64));
  for (int i = 0; i < 3; ++i) out.push_back(at::empty({n}, i32));
  out.push_back(at::empty({n, Dz}, sb.options()));
  for (int i = 0; i < 4; ++i) out.push_back(at::empty({n, Dy, Dx}, sb.options()));
  out.push_back(at::empty({n, Dy, Dx}, sb.options().dtype(at::kByte)));
  xpg_pass1_out o{};
  double **d_out[14] = {&o.bg, &o.nbg, &o.isum, &o.vsum, &o.vsum0, &o.xo, &o.yo, &o.zo, &o.m_d, &o.m_m, &o.frac_rec,
                        &o.nfg, &o.zeta, &o.seff};
  for (int i = 0; i < 14; ++i) *d_out[i] = out[i].data_ptr<double>();
  o.n_over = out[14].data_ptr<int32_t>();
  o.n_fg_masked = out[15].data_ptr<int32_t>();
  o.n_fg_nb = out[16].data_ptr<int32_t>();
  o.fz = out[17].data_ptr<float>();
  o.img2 = out[18].data_ptr<float>();
  o.J = out[19].data_ptr<float>();
  o.u1 = out[20].data_ptr<float>();
  o.u2 = out[21].data_ptr<float>();
  o.all_valid2 = out[22].data_ptr<uint8_t>();
  xpg_profile_acc acc{};
  TORCH_CHECK(acc_num.has_value() == acc_den.has_value(), "acc_num and acc_den go together");
  if (acc_num.has_value()) {
    TORCH_CHECK(acc_num->dim() == 3 && acc_num->sizes() == acc_den->sizes(), "profile grids: (n_regions, 41, 41)");
    acc.num = const_cast<int64_t *>(dptr<int64_t>(*acc_num, at::kLong, "acc_num"));
    acc.den = const_cast<int64_t *>(dptr<int64_t>(*acc_den, at::kLong, "acc_den"));
    acc.n_regions = (int32_t)acc_num->size(0);
    acc.copies = 1;
  }
  check_rc(xpg_pass1(&d, &k, &R.r, sb.data_ptr<float>(), &o, acc_num.has_value() ? &acc : nullptr, stream()), "pass1");
  return out;
}

std::tuple<Tensor, Tensor> accumulate_profiles(const Tensor &img2, const Tensor &ok2, const Tensor &J,
                                               const Tensor &u1, const Tensor &u2, const Tensor &region,
                                               const Tensor &isum, const Tensor &sel, const Tensor &seff,
                                               int64_t n_regions) {
  TORCH_CHECK(img2.dim() == 3, "img2: (n, Dy, Dx)");
  const int64_t n = img2.size(0), P = img2.size(1) * img2.size(2);
  const auto i64 = img2.options().dtype(at::kLong);
  Tensor num = at::zeros({n_regions, XPG_PROFILE_GRID, XPG_PROFILE_GRID}, i64);
  Tensor den = at::zeros({n_regions, XPG_PROFILE_GRID, XPG_PROFILE_GRID}, i64);
  xpg_profile_acc acc{num.data_ptr<int64_t>(), den.data_ptr<int64_t>(), (int32_t)n_regions, 1};
  check_rc(xpg_accumulate_profiles(dptr<float>(img2, at::kFloat, "img2"), dptr<uint8_t>(ok2, at::kByte, "ok2", n * P),
                                   dptr<float>(J, at::kFloat, "J", n * P), dptr<float>(u1, at::kFloat, "u1", n * P),
                                   dptr<float>(u2, at::kFloat, "u2", n * P),
                                   dptr<int32_t>(region, at::kInt, "region", n), dptr<double>(isum, at::kDouble, "isum", n),
                                   dptr<uint8_t>(sel, at::kByte, "sel", n), dptr<double>(seff, at::kDouble, "seff", n),
                                   (int32_t)n, (int32_t)img2.size(1), (int32_t)img2.size(2), &acc, stream()),
           "accumulate_profiles");
  return {num.to(at::kDouble) / FX_SCALE, den.to(at::kDouble) / FX_SCALE};
}

std::tuple<Tensor, Tensor, Tensor> normalise_profiles(const Tensor &num_fx, const Tensor &den_fx, const Tensor &gauss) {
  TORCH_CHECK(num_fx.dim() == 3 && num_fx.sizes() == den_fx.sizes(), "profile grids: (n_regions, 41, 41)");
  const int64_t nr = num_fx.size(0);
  xpg_profile_acc acc{const_cast<int64_t *>(dptr<int64_t>(num_fx, at::kLong, "num_fx")),
                      const_cast<int64_t *>(dptr<int64_t>(den_fx, at::kLong, "den_fx")), (int32_t)nr, 1};
  Tensor prof = at::empty(num_fx.sizes(), num_fx.options().dtype(at::kFloat));
  Tensor n64 = at::empty(num_fx.sizes(), num_fx.options().dtype(at::kDouble));
  Tensor d64 = at::empty(num_fx.sizes(), num_fx.options().dtype(at::kDouble));
  check_rc(xpg_normalise_profiles(&acc, dptr<float>(gauss, at::kFloat, "gauss", XPG_PROFILE_GRID * XPG_PROFILE_GRID),
                                  prof.data_ptr<float>(), n64.data_ptr<double>(), d64.data_ptr<double>(), stream()),
           "normalise_profiles");
  return {prof, n64, d64};
}

std::vector<Tensor> pass2(const Tensor &sb, const c10::List<OptT> &rows, const c10::List<OptT> &det,
                          at::ArrayRef<double> det_scalars, at::ArrayRef<double> call,
                          c10::optional<at::ArrayRef<double>> rcw, const Tensor &prof, const Tensor &bg, const Tensor &nbg) {
  check_sb(sb);
  const int64_t n = sb.size(0);
  Rows R = make_refl(rows, n, sb.size(1), sb.size(2), sb.size(3));
  const xpg_detector d = make_det(det, det_scalars);
  const xpg_call k = make_call(call, rcw);
  TORCH_CHECK(prof.dim() == 3, "prof: (n_regions, 41, 41)");
  std::vector<Tensor> out;
  for (int i = 0; i < 9; ++i) out.push_back(at::empty({n}, sb.options().dtype(at::kDouble)));
  xpg_pass2_out o{};
  double **d_out[9] = {&o.iprf, &o.vprf, &o.vprf0, &o.corr, &o.p_valid, &o.p_total, &o.ipartner, &o.rho, &o.p_fg};
  for (int i = 0; i < 9; ++i) *d_out[i] = out[i].data_ptr<double>();
  check_rc(xpg_pass2(&d, &k, &R.r, sb.data_ptr<float>(), dptr<float>(prof, at::kFloat, "prof"), (int32_t)prof.size(0),
                     dptr<double>(bg, at::kDouble, "bg", n), dptr<double>(nbg, at::kDouble, "nbg", n), 0, &o, stream()),
           "pass2");
  return out;
}

}  // namespace

TORCH_LIBRARY(xint, m) {
  m.def("extract(Tensor frames, int frame0, int min_valid, int flag_value, int overload, bool use_flag_value, "
        "Tensor?[] rows, int[] box) -> Tensor");
  m.def("pass1(Tensor sb, Tensor?[] rows, Tensor?[] det, float[] det_scalars, float[] call, float[]? rcw, "
        "Tensor? acc_num=None, Tensor? acc_den=None) -> Tensor[]");
  m.def("accumulate_profiles(Tensor img2, Tensor ok2, Tensor J, Tensor u1, Tensor u2, Tensor region, Tensor isum, "
        "Tensor sel, Tensor seff, int n_regions) -> (Tensor, Tensor)");
  m.def("normalise_profiles(Tensor num_fx, Tensor den_fx, Tensor gauss) -> (Tensor, Tensor, Tensor)");
  m.def("pass2(Tensor sb, Tensor?[] rows, Tensor?[] det, float[] det_scalars, float[] call, float[]? rcw, "
        "Tensor prof, Tensor bg, Tensor nbg) -> Tensor[]");
}

TORCH_LIBRARY_IMPL(xint, CUDA, m) {
  m.impl("extract", &extract);
  m.impl("pass1", &pass1);
  m.impl("accumulate_profiles", &accumulate_profiles);
  m.impl("normalise_profiles", &normalise_profiles);
  m.impl("pass2", &pass2);
}
