// cxx_abi.cpp -- C++-linkage copies of the single-point eval functions
// (1-D, 2-D and 3-D).
//
// The reference declares eval_multi_UBspline_3d_* without extern "C"
// (src/multi_bspline_eval_d.h:31-110) and compiles each implementation a
// second time as C++ (src/multi_bspline_eval_std_d_cpp.cc:21), so C++
// objects built against its headers reference C++-mangled names.  These
// forwarders export those names; the C symbols (include/*.h) do the work.
#include <complex>

#include "multi_bspline_structs.h"

// The C implementations, reached through their unmangled symbol names.

void c_eval_s(multi_UBspline_3d_s *, double, double, double, float *) __asm__("eval_multi_UBspline_3d_s");
void c_eval_s_vg(multi_UBspline_3d_s *, double, double, double, float *, float *) __asm__("eval_multi_UBspline_3d_s_vg");
void c_eval_s_vgl(multi_UBspline_3d_s *, double, double, double, float *, float *, float *) __asm__("eval_multi_UBspline_3d_s_vgl");
void c_eval_s_vgh(multi_UBspline_3d_s *, double, double, double, float *, float *, float *) __asm__("eval_multi_UBspline_3d_s_vgh");
void c_eval_d(multi_UBspline_3d_d *, double, double, double, double *) __asm__("eval_multi_UBspline_3d_d");
void c_eval_d_vg(multi_UBspline_3d_d *, double, double, double, double *, double *) __asm__("eval_multi_UBspline_3d_d_vg");
void c_eval_d_vgl(multi_UBspline_3d_d *, double, double, double, double *, double *, double *) __asm__("eval_multi_UBspline_3d_d_vgl");
void c_eval_d_vgh(multi_UBspline_3d_d *, double, double, double, double *, double *, double *) __asm__("eval_multi_UBspline_3d_d_vgh");
void c_eval_c(multi_UBspline_3d_c *, double, double, double, std::complex<float> *) __asm__("eval_multi_UBspline_3d_c");
void c_eval_c_vg(multi_UBspline_3d_c *, double, double, double, std::complex<float> *, std::complex<float> *) __asm__("eval_multi_UBspline_3d_c_vg");
void c_eval_c_vgl(multi_UBspline_3d_c *, double, double, double, std::complex<float> *, std::complex<float> *, std::complex<float> *) __asm__("eval_multi_UBspline_3d_c_vgl");
void c_eval_c_vgh(multi_UBspline_3d_c *, double, double, double, std::complex<float> *, std::complex<float> *, std::complex<float> *) __asm__("eval_multi_UBspline_3d_c_vgh");
void c_eval_z(multi_UBspline_3d_z *, double, double, double, std::complex<double> *) __asm__("eval_multi_UBspline_3d_z");
void c_eval_z_vg(multi_UBspline_3d_z *, double, double, double, std::complex<double> *, std::complex<double> *) __asm__("eval_multi_UBspline_3d_z_vg");
void c_eval_z_vgl(multi_UBspline_3d_z *, double, double, double, std::complex<double> *, std::complex<double> *, std::complex<double> *) __asm__("eval_multi_UBspline_3d_z_vgl");
void c_eval_z_vgh(multi_UBspline_3d_z *, double, double, double, std::complex<double> *, std::complex<double> *, std::complex<double> *) __asm__("eval_multi_UBspline_3d_z_vgh");

void eval_multi_UBspline_3d_s(multi_UBspline_3d_s *s, double x, double y, double z, float *v) { c_eval_s(s, x, y, z, v); }
void eval_multi_UBspline_3d_s_vg(multi_UBspline_3d_s *s, double x, double y, double z, float *v, float *g) {
  c_eval_s_vg(s, x, y, z, v, g);
}
void eval_multi_UBspline_3d_s_vgl(multi_UBspline_3d_s *s, double x, double y, double z, float *v, float *g,
                                 float *l) {
  c_eval_s_vgl(s, x, y, z, v, g, l);
}
void eval_multi_UBspline_3d_s_vgh(multi_UBspline_3d_s *s, double x, double y, double z, float *v, float *g,
                                 float *h) {
  c_eval_s_vgh(s, x, y, z, v, g, h);
}
void eval_multi_UBspline_3d_d(multi_UBspline_3d_d *s, double x, double y, double z, double *v) { c_eval_d(s, x, y, z, v); }
void eval_multi_UBspline_3d_d_vg(multi_UBspline_3d_d *s, double x, double y, double z, double *v, double *g) {
  c_eval_d_vg(s, x, y, z, v, g);
}
void eval_multi_UBspline_3d_d_vgl(multi_UBspline_3d_d *s, double x, double y, double z, double *v, double *g,
                                 double *l) {
  c_eval_d_vgl(s, x, y, z, v, g, l);
}
void eval_multi_UBspline_3d_d_vgh(multi_UBspline_3d_d *s, double x, double y, double z, double *v, double *g,
                                 double *h) {
  c_eval_d_vgh(s, x, y, z, v, g, h);
}
void eval_multi_UBspline_3d_c(multi_UBspline_3d_c *s, double x, double y, double z, std::complex<float> *v) { c_eval_c(s, x, y, z, v); }
void eval_multi_UBspline_3d_c_vg(multi_UBspline_3d_c *s, double x, double y, double z, std::complex<float> *v, std::complex<float> *g) {
  c_eval_c_vg(s, x, y, z, v, g);
}
void eval_multi_UBspline_3d_c_vgl(multi_UBspline_3d_c *s, double x, double y, double z, std::complex<float> *v, std::complex<float> *g,
                                 std::complex<float> *l) {
  c_eval_c_vgl(s, x, y, z, v, g, l);
}
void eval_multi_UBspline_3d_c_vgh(multi_UBspline_3d_c *s, double x, double y, double z, std::complex<float> *v, std::complex<float> *g,
                                 std::complex<float> *h) {
  c_eval_c_vgh(s, x, y, z, v, g, h);
}
void eval_multi_UBspline_3d_z(multi_UBspline_3d_z *s, double x, double y, double z, std::complex<double> *v) { c_eval_z(s, x, y, z, v); }
void eval_multi_UBspline_3d_z_vg(multi_UBspline_3d_z *s, double x, double y, double z, std::complex<double> *v, std::complex<double> *g) {
  c_eval_z_vg(s, x, y, z, v, g);
}
void eval_multi_UBspline_3d_z_vgl(multi_UBspline_3d_z *s, double x, double y, double z, std::complex<double> *v, std::complex<double> *g,
                                 std::complex<double> *l) {
  c_eval_z_vgl(s, x, y, z, v, g, l);
}
void eval_multi_UBspline_3d_z_vgh(multi_UBspline_3d_z *s, double x, double y, double z, std::complex<double> *v, std::complex<double> *g,
                                 std::complex<double> *h) {
  c_eval_z_vgh(s, x, y, z, v, g, h);
}

// vghgh (src/multi_bspline_eval_std_d_impl.h:868 and the s/c/z files)
#define ESK_VGHGH_FWD(T, TY)                                                                        \
  void c_eval_##T##_vghgh(multi_UBspline_3d_##T *, double, double, double, TY *, TY *, TY *, TY *)  \
      __asm__("eval_multi_UBspline_3d_" #T "_vghgh");                                              \
  void eval_multi_UBspline_3d_##T##_vghgh(multi_UBspline_3d_##T *s, double x, double y, double z,   \
                                          TY *v, TY *g, TY *h, TY *gh) {                            \
    c_eval_##T##_vghgh(s, x, y, z, v, g, h, gh);                                                    \
  }
ESK_VGHGH_FWD(s, float)
ESK_VGHGH_FWD(d, double)
ESK_VGHGH_FWD(c, std::complex<float>)
ESK_VGHGH_FWD(z, std::complex<double>)

// 1-D and 2-D (src/multi_bspline_eval_std_d_impl.h:37-476 and the s/c/z files,
// compiled as C++ by the same *_cpp.cc units)
#define ESK_FWD12(T, TY)                                                                            \
  void c_eval_1d_##T(multi_UBspline_1d_##T *, double, TY *) __asm__("eval_multi_UBspline_1d_" #T);  \
  void c_eval_1d_##T##_vg(multi_UBspline_1d_##T *, double, TY *, TY *)                              \
      __asm__("eval_multi_UBspline_1d_" #T "_vg");                                                  \
  void c_eval_1d_##T##_vgl(multi_UBspline_1d_##T *, double, TY *, TY *, TY *)                       \
      __asm__("eval_multi_UBspline_1d_" #T "_vgl");                                                 \
  void c_eval_1d_##T##_vgh(multi_UBspline_1d_##T *, double, TY *, TY *, TY *)                       \
      __asm__("eval_multi_UBspline_1d_" #T "_vgh");                                                 \
  void c_eval_2d_##T(multi_UBspline_2d_##T *, double, double, TY *) __asm__("eval_multi_UBspline_2d_" #T); \
  void c_eval_2d_##T##_vg(multi_UBspline_2d_##T *, double, double, TY *, TY *)                      \
      __asm__("eval_multi_UBspline_2d_" #T "_vg");                                                  \
  void c_eval_2d_##T##_vgl(multi_UBspline_2d_##T *, double, double, TY *, TY *, TY *)               \
      __asm__("eval_multi_UBspline_2d_" #T "_vgl");                                                 \
  void c_eval_2d_##T##_vgh(multi_UBspline_2d_##T *, double, double, TY *, TY *, TY *)               \
      __asm__("eval_multi_UBspline_2d_" #T "_vgh");                                                 \
  void eval_multi_UBspline_1d_##T(multi_UBspline_1d_##T *s, double x, TY *v) { c_eval_1d_##T(s, x, v); } \
  void eval_multi_UBspline_1d_##T##_vg(multi_UBspline_1d_##T *s, double x, TY *v, TY *g) {          \
    c_eval_1d_##T##_vg(s, x, v, g);                                                                 \
  }                                                                                                 \
  void eval_multi_UBspline_1d_##T##_vgl(multi_UBspline_1d_##T *s, double x, TY *v, TY *g, TY *l) {  \
    c_eval_1d_##T##_vgl(s, x, v, g, l);                                                             \
  }                                                                                                 \
  void eval_multi_UBspline_1d_##T##_vgh(multi_UBspline_1d_##T *s, double x, TY *v, TY *g, TY *h) {  \
    c_eval_1d_##T##_vgh(s, x, v, g, h);                                                             \
  }                                                                                                 \
  void eval_multi_UBspline_2d_##T(multi_UBspline_2d_##T *s, double x, double y, TY *v) {            \
    c_eval_2d_##T(s, x, y, v);                                                                      \
  }                                                                                                 \
  void eval_multi_UBspline_2d_##T##_vg(multi_UBspline_2d_##T *s, double x, double y, TY *v, TY *g) { \
    c_eval_2d_##T##_vg(s, x, y, v, g);                                                              \
  }                                                                                                 \
  void eval_multi_UBspline_2d_##T##_vgl(multi_UBspline_2d_##T *s, double x, double y, TY *v, TY *g, \
                                        TY *l) {                                                    \
    c_eval_2d_##T##_vgl(s, x, y, v, g, l);                                                          \
  }                                                                                                 \
  void eval_multi_UBspline_2d_##T##_vgh(multi_UBspline_2d_##T *s, double x, double y, TY *v, TY *g, \
                                        TY *h) {                                                    \
    c_eval_2d_##T##_vgh(s, x, y, v, g, h);                                                          \
  }
ESK_FWD12(s, float)
ESK_FWD12(d, double)
ESK_FWD12(c, std::complex<float>)
ESK_FWD12(z, std::complex<double>)
