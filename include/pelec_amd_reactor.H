// pelec_amd_reactor.H -- C++ shim with the method names of pele::physics::reactions::ReactorBase,
// forwarding to the C-ABI of pelec_amd.h. A PeleC build registers this class where it would create
// a PelePhysics reactor (Source/PeleC.cpp:2032-2047) and calls react() at the one call site,
// Source/React.cpp:222-228; nothing else in PeleC changes.
//
// The ReactorBase header lives in the absent PelePhysics submodule, so the shim does not derive
// from it. It is templated on the caller's Box / Array4 types instead: any Box with smallEnd(d) /
// bigEnd(d) and any Array4 with the public members p, jstride, kstride, nstride, ncomp and
// begin / end of a Dim3-like type with members x, y, z (amrex::Box and amrex::Array4 have exactly
// these) binds without this header including AMReX.
#ifndef PELEC_AMD_REACTOR_H
#define PELEC_AMD_REACTOR_H

#include <stdexcept>
#include <string>
#include <vector>

#include "pelec_amd.h"

namespace pelec_amd {

class Reactor {
 public:
  // ReactorBase::create(chem_integrator): "ReactorCvode", "ReactorRK64" or "ReactorNull".
  static Reactor* create(const std::string& chem_integrator) {
    prc_reactor* h = nullptr;
    check(prc_create(chem_integrator.c_str(), &h), "prc_create");
    return new Reactor(h);
  }

  ~Reactor() {
    if (h_) (void)prc_destroy(h_);
  }
  Reactor(const Reactor&) = delete;
  Reactor& operator=(const Reactor&) = delete;

  // ReactorBase::init(reactor_type, ncells); PeleC calls init(1, 1) (Source/PeleC.cpp:2040).
  int init(int reactor_type, int ncells) { return prc_init(h_, reactor_type, ncells); }

  // ReactorBase::set_typ_vals_ode (Source/React.cpp:11,25): NUM_SPECIES + 1 values.
  void set_typ_vals_ode(const std::vector<double>& typ) {
    check(prc_set_typ_vals_ode(h_, typ.data(), static_cast<int>(typ.size())), "set_typ_vals_ode");
  }

  // ReactorBase::close() (Source/PeleC.cpp:2046).
  void close() { (void)prc_close(h_); }

  // Knobs the reference reads from ParmParse (ode.rtol, ode.atol, ode.max_steps,
  // cvode.max_order) and the gfx950 form selection (prc.form, prc.lu_fp32).
  void set(const char* key, double v) { check(prc_set_param_real(h_, key, v), key); }
  void set(const char* key, int v) { check(prc_set_param_int(h_, key, v), key); }

  // ReactorBase::react(box, rY, rYsrc_ext, T, rEner, rEner_src_ext, FC, mask, dt_react, time,
  // stream) as called at Source/React.cpp:222-228. Stream-ordered; the caller synchronises as
  // PeleC does (React.cpp:230). Returns PRC_OK or a negative PRC_E* code.
  template <class Box, class RA4, class IA4>
  int react(const Box& box, const RA4& rY, const RA4& rYsrc_ext, const RA4& T, const RA4& rEner,
            const RA4& rEner_src_ext, const RA4& FC, const IA4& mask, double& dt_react,
            double& time, void* stream = nullptr) {
    int32_t lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = box.smallEnd(d);
      hi[d] = box.bigEnd(d);
    }
    return prc_react_box(h_, lo, hi, view(rY), view(rYsrc_ext), view(T), view(rEner),
                         view(rEner_src_ext), view(FC), iview(mask), &dt_react, &time, stream);
  }

  // Cells of the previous react call whose integration failed (synchronises the stream).
  int64_t last_failures(void* stream = nullptr) { return prc_last_failures(h_, stream); }

  prc_reactor* handle() const { return h_; }

 private:
  explicit Reactor(prc_reactor* h) : h_(h) {}

  static void check(int rc, const char* what) {
    if (rc != PRC_OK) throw std::runtime_error(std::string("pelec_amd: ") + what + " failed");
  }

  template <class D3>
  static int32_t comp(const D3& v, int d) {
    return static_cast<int32_t>(d == 0 ? v.x : (d == 1 ? v.y : v.z));
  }

  template <class A4>
  static prc_array4 view(const A4& a) {
    prc_array4 v;
    v.p = const_cast<double*>(static_cast<const double*>(a.p));
    v.jstride = a.jstride;
    v.kstride = a.kstride;
    v.nstride = a.nstride;
    for (int d = 0; d < 3; ++d) {
      v.begin[d] = comp(a.begin, d);
      v.end[d] = comp(a.end, d);
    }
    v.ncomp = static_cast<int32_t>(a.ncomp);
    return v;
  }

  template <class A4>
  static prc_iarray4 iview(const A4& a) {
    prc_iarray4 v;
    v.p = const_cast<int32_t*>(static_cast<const int32_t*>(a.p));
    v.jstride = a.jstride;
    v.kstride = a.kstride;
    v.nstride = a.nstride;
    for (int d = 0; d < 3; ++d) {
      v.begin[d] = comp(a.begin, d);
      v.end[d] = comp(a.end, d);
    }
    v.ncomp = static_cast<int32_t>(a.ncomp);
    return v;
  }

  prc_reactor* h_;
};

}  // namespace pelec_amd

#endif  // PELEC_AMD_REACTOR_H
