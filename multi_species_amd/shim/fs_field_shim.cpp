// The reference-side binding of libfarrsight_hip: ElectricField subclasses
// (inc/FieldStructure.hpp:87-93) whose operator() runs the HIP solvers, built
// against the reference's own header so that FARRSIGHT's AMRSimulation calls
// them through its existing virtual interface (AMRSimulation_field.cpp:68-69
// and :167-168) with no other change. The base class's out-of-line members
// (ElectricField::~ElectricField, its vtable and typeinfo) come from the
// reference's FieldStructure.cpp, which the program that uses this binding
// contains; this shared object leaves them to the dynamic linker.
//
//   E_MQ_DirectSum_HIP          replaces E_MQ_DirectSum          (FieldStructure.cpp:106-156)
//   E_MQ_DirectSum_openbcs_HIP  replaces E_MQ_DirectSum_openbcs  (FieldStructure.cpp:166-191)
//   E_MQ_Treecode_HIP           replaces E_MQ_Treecode           (FieldStructure.cpp:230-382)
//
// Same contract as the classes they replace: es[0:nt] overwritten, the inputs
// untouched, synchronous. A failure throws std::runtime_error with the
// library's message (the reference has no error channel).
//
// The extern "C" functions at the end make and call these objects through an
// ElectricField pointer, for callers outside C++ (the tests).
#include <iostream>
#include <stdexcept>
#include <string>

#include "FieldStructure.hpp"
#include "farrsight_hip.h"

namespace {

fs_ctx* make_ctx(int device) {
    fs_ctx* ctx = nullptr;
    if (fs_ctx_create(&ctx, device) != FS_OK) throw std::runtime_error(fs_last_error());
    return ctx;
}

void check(int rc) {
    if (rc != FS_OK) throw std::runtime_error(fs_last_error());
}

}  // namespace

class E_MQ_DirectSum_HIP : public ElectricField {
  public:
    E_MQ_DirectSum_HIP(double L, double epsilon, int device = 0)
        : ctx(make_ctx(device)), L(L), epsilon(epsilon) {}
    ~E_MQ_DirectSum_HIP() override { fs_ctx_destroy(ctx); }
    void operator()(double* es, double* targets, int nt, double* sources, double* q_ws,
                    int ns) override {
        check(fs_emq_periodic_f64(ctx, es, targets, nt, sources, q_ws, ns, L, epsilon));
    }
    void print_field_obj() override {
        std::cout << "MQ kernel, direct sum on the GPU (HIP, gfx950)" << std::endl;
        std::cout << "epsilon = " << epsilon << ", periodic, L = " << L << std::endl;
    }

  private:
    fs_ctx* ctx;
    double L, epsilon;
};

class E_MQ_DirectSum_openbcs_HIP : public ElectricField {
  public:
    explicit E_MQ_DirectSum_openbcs_HIP(double epsilon, int device = 0)
        : ctx(make_ctx(device)), epsilon(epsilon) {}
    ~E_MQ_DirectSum_openbcs_HIP() override { fs_ctx_destroy(ctx); }
    void operator()(double* es, double* targets, int nt, double* sources, double* q_ws,
                    int ns) override {
        check(fs_emq_open_f64(ctx, es, targets, nt, sources, q_ws, ns, epsilon));
    }
    void print_field_obj() override {
        std::cout << "MQ kernel, direct sum on the GPU (HIP, gfx950)" << std::endl;
        std::cout << "epsilon = " << epsilon << ", open boundaries" << std::endl;
    }

  private:
    fs_ctx* ctx;
    double epsilon;
};

class E_MQ_Treecode_HIP : public ElectricField {
  public:
    E_MQ_Treecode_HIP(double L, double epsilon, double mac, int degree, int max_source,
                      int device = 0)
        : ctx(make_ctx(device)), L(L), epsilon(epsilon), mac(mac), degree(degree),
          max_source(max_source) {}
    ~E_MQ_Treecode_HIP() override { fs_ctx_destroy(ctx); }
    // like E_MQ_Treecode, evaluates at the sources (targets and nt unused)
    void operator()(double* es, double* targets, int nt, double* sources, double* q_ws,
                    int ns) override {
        (void)targets;
        (void)nt;
        check(fs_emq_treecode_f64(ctx, es, sources, q_ws, ns, L, epsilon, mac, degree,
                                  max_source, nullptr));
    }
    void print_field_obj() override {
        std::cout << "MQ kernel, treecode on the GPU (HIP, gfx950)" << std::endl;
        std::cout << "epsilon = " << epsilon << ", periodic, L = " << L << ", mac = " << mac
                  << ", degree = " << degree << ", max_source = " << max_source << std::endl;
    }

  private:
    fs_ctx* ctx;
    double L, epsilon, mac;
    int degree, max_source;
};

extern "C" {

// kind 0: E_MQ_DirectSum_HIP(L, eps), 1: E_MQ_DirectSum_openbcs_HIP(eps),
// 2: E_MQ_Treecode_HIP(L, eps, mac, degree, max_source); 10 and 11 the
// reference's own E_MQ_DirectSum and E_MQ_DirectSum_openbcs. NULL on failure.
ElectricField* fs_shim_new_field(int kind, double L, double eps, double mac, int degree,
                                 int max_source) {
    try {
        switch (kind) {
            case 0: return new E_MQ_DirectSum_HIP(L, eps);
            case 1: return new E_MQ_DirectSum_openbcs_HIP(eps);
            case 2: return new E_MQ_Treecode_HIP(L, eps, mac, degree, max_source);
            case 10: return new E_MQ_DirectSum(L, eps);
            case 11: return new E_MQ_DirectSum_openbcs(eps);
            default: return nullptr;
        }
    } catch (const std::exception&) {
        return nullptr;
    }
}

// field->operator()(...) through the base class; 0, or -1 if it threw.
int fs_shim_call(ElectricField* field, double* es, double* targets, int nt, double* sources,
                 double* q_ws, int ns) {
    try {
        (*field)(es, targets, nt, sources, q_ws, ns);
    } catch (const std::exception& e) {
        std::cerr << "ElectricField::operator() failed: " << e.what() << std::endl;
        return -1;
    }
    return 0;
}

void fs_shim_print(ElectricField* field) { field->print_field_obj(); }

void fs_shim_delete(ElectricField* field) { delete field; }

}  // extern "C"
