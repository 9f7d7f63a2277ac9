// A C++ caller of include/bmagwa_hip.h shaped like the reference's use of
// Model (src/model.hpp:49-392) from its Sampler: a HipModel class owns one
// bmg_model and forwards the Model methods; the program reads a dataset and a
// list of operations, applies them and prints every result.
//
// Input file (little endian): uint32 n, uint64 p, p*ceil(n/4) packed bytes,
// n doubles y, then a bmg_opts struct, uint32 count, and count operations of
// int32 op, uint32 arg, double d: op 0 add_term(arg, type 0, d), 1
// remove_term(arg), 2 sample_beta_sigma2, 3 compute_pve + RB scan.
// Output (stdout): one line per operation.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "bmagwa_hip.h"

static void check(int rc) {
  if (rc != BMG_OK) throw std::runtime_error(bmg_last_error());
}

class HipModel {
 public:
  HipModel(bmg_ctx* ctx, const bmg_opts& opts) { check(bmg_model_create_sampler(ctx, &opts, &m_)); }
  ~HipModel() { bmg_model_destroy(m_); }
  HipModel(const HipModel&) = delete;
  HipModel& operator=(const HipModel& src) { check(bmg_model_copy(m_, src.m_)); return *this; }

  void add_term(size_t ind, char t_ind, const double* inv_tau2_alpha2) {
    check(bmg_model_add_term(m_, (uint32_t)ind, t_ind, inv_tau2_alpha2));
  }
  void remove_term(size_t model_ind) { check(bmg_model_remove_term(m_, (int)model_ind)); }
  double log_likelihood() const { double v; check(bmg_model_log_likelihood(m_, &v)); return v; }
  size_t size() const { int k; check(bmg_model_size(m_, &k)); return (size_t)k; }
  void sample_beta_sigma2() { check(bmg_model_sample_beta_sigma2(m_)); }
  void compute_pve(double* pves) { check(bmg_model_compute_pve(m_, pves)); }
  std::vector<double> beta(double* sigma2, std::vector<double>* y_hat, size_t n) const {
    int len = 0;
    check(bmg_model_beta(m_, nullptr, nullptr, nullptr, &len));
    std::vector<double> b(len);
    y_hat->resize(n);
    check(bmg_model_beta(m_, b.data(), sigma2, y_hat->data(), &len));
    return b;
  }
  void p_raoblackwell(const double* y_hat, double* p_r) { check(bmg_model_rb(m_, y_hat, p_r)); }

 private:
  bmg_model* m_ = nullptr;
};

template <typename T>
static T get(FILE* f) {
  T v;
  if (fread(&v, sizeof(T), 1, f) != 1) throw std::runtime_error("short input");
  return v;
}

int main(int argc, char** argv) {
  try {
    if (argc != 2) throw std::runtime_error("usage: model_caller <input>");
    FILE* f = fopen(argv[1], "rb");
    if (!f) throw std::runtime_error("cannot open input");
    const uint32_t n = get<uint32_t>(f);
    const uint64_t p = get<uint64_t>(f);
    std::vector<uint8_t> body((size_t)p * ((n + 3) / 4));
    std::vector<double> y(n);
    if (fread(body.data(), 1, body.size(), f) != body.size() || fread(y.data(), 8, n, f) != n)
      throw std::runtime_error("short input");
    const bmg_opts opts = get<bmg_opts>(f);
    const uint32_t count = get<uint32_t>(f);
    int32_t types[1] = {BMG_A};
    bmg_ctx* ctx = nullptr;
    check(bmg_ctx_create(0, body.data(), n, p, y.data(), types, 1, 0, &ctx));
    {
      HipModel current(ctx, opts), proposed(ctx, opts);
      std::vector<double> p_r(p), y_hat;
      for (uint32_t i = 0; i < count; ++i) {
        const int32_t op = get<int32_t>(f);
        const uint32_t arg = get<uint32_t>(f);
        const double d = get<double>(f);
        if (op == 0) {
          const double dd[2] = {d, d};
          proposed.add_term(arg, 0, dd);
          current = proposed;
          printf("add %u %zu %.17g\n", arg, current.size(), current.log_likelihood());
        } else if (op == 1) {
          proposed.remove_term(arg);
          current = proposed;
          printf("remove %u %zu %.17g\n", arg, current.size(), current.log_likelihood());
        } else if (op == 2) {
          current.sample_beta_sigma2();
          double s2 = 0;
          std::vector<double> b = current.beta(&s2, &y_hat, n);
          printf("beta %.17g", s2);
          for (double v : b) printf(" %.17g", v);
          printf("\n");
        } else {
          double pves[3];
          current.compute_pve(pves);
          double s2 = 0;
          current.beta(&s2, &y_hat, n);
          current.p_raoblackwell(y_hat.data(), p_r.data());
          printf("pve %.17g %.17g %.17g\nrb", pves[0], pves[1], pves[2]);
          for (double v : p_r) printf(" %.17g", v);
          printf("\n");
        }
      }
    }
    fclose(f);
    check(bmg_ctx_destroy(ctx));
    return 0;
  } catch (const std::exception& e) {
    fprintf(stderr, "model_caller: %s\n", e.what());
    return 1;
  }
}
