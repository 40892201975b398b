/*
 * pittpack_hip.hpp -- C++ class over the C ABI with the public set-up / solve methods of the
 * reference's PencilDcmp (src/include/pencilDcmp.hpp:441-466), so a caller written against
 * make_Poisson (src/pencilDcmp.cpp:6432-6440) and main.cpp's call order (src/main.cpp:73-118)
 * switches by changing the type:
 *
 *     auto M = make_Poisson(argc, argv, Nx, Ny, Nz);
 *     char bc[6] = {'N', 'N', 'N', 'N', 'D', 'D'};  M->assignBoundary(bc);
 *     double X[6] = {0, 1, 0, 1, 0, 1};            M->setBox(X);
 *     M->setCoords(2);
 *     M->assignRhs(rhs);                            // or M->initializeTrigonometric();
 *     M->pittPack();
 *     double err = M->getError();  M->getSolution(u);  M->IO(1, 2, 0);
 *
 * Failures throw PittPackException carrying the C ABI's code (the reference's numbers 0..23
 * plus the codes >= 24 of pittpack_hip.h) instead of calling exit(1).  Header only; link
 * with -lpittpack_hip.
 */
#ifndef PITTPACK_HIP_HPP
#define PITTPACK_HIP_HPP

#include <memory>
#include <stdexcept>
#include <string>

#include "pittpack_hip.h"

class PittPackException : public std::runtime_error {
  public:
    PittPackException(int code, const char *where)
        : std::runtime_error(std::string(where) + ": " + pp_get_error_enum(code)), code_(code)
    {
    }
    int code() const { return code_; }

  private:
    int code_;
};

class PoissonHIP {
  public:
    PoissonHIP(int nx, int ny, int nz) { check(pp_create(nx, ny, nz, &h_), "PoissonHIP"); }
    ~PoissonHIP()
    {
        if (h_) pp_destroy(h_);
    }
    PoissonHIP(const PoissonHIP &) = delete;
    PoissonHIP &operator=(const PoissonHIP &) = delete;

    // PencilDcmp::assignBoundary(char*), setBox(double*), setCoords(int), assignRhs(double*),
    // pittPack(), getError(), initializeTrigonometric(), fillTrigonometric(double*), IO(...)
    void assignBoundary(const char *bc) { check(pp_assign_boundary(h_, bc), "assignBoundary"); }
    void setBox(const double *X) { check(pp_set_box(h_, X), "setBox"); }
    void setCoords(int dir) { check(pp_set_coords(h_, dir), "setCoords"); }
    void assignRhs(const double *rhs) { check(pp_assign_rhs(h_, rhs), "assignRhs"); }
    void pittPack() { check(pp_pittpack(h_), "pittPack"); }
    // part 0 / 1: real / imaginary part of the reference's complex manufactured field
    void initializeTrigonometric(int part = 0) { check(pp_init_trigonometric(h_, part), "initializeTrigonometric"); }
    void fillTrigonometric(double *rhs) { check(pp_fill_trigonometric(h_, rhs), "fillTrigonometric"); }
    double getError()
    {
        double e = 0.0;
        check(pp_get_error(h_, &e), "getError");
        return e;
    }
    // dir / aligndir select the reference's chunk orientation; one device holds the whole grid
    // in the orientation the reference writes with (dir 2, aligndir 0), so they are accepted
    // and have no effect.  Files go to soln/ as in the reference.
    void IO(int app, int dir = 2, int aligndir = 0, const char *folder = "soln")
    {
        (void)dir;
        (void)aligndir;
        check(pp_io(h_, folder, app), "IO");
    }

    // no counterpart in the reference (P is protected there)
    void getSolution(double *u) { check(pp_get_solution(h_, u), "getSolution"); }
    void solveDevice(double *d_field) { check(pp_solve_device(h_, d_field), "solveDevice"); }
    void solveDevice(const double *d_rhs, double *d_u) { check(pp_solve_device_oop(h_, d_rhs, d_u), "solveDevice"); }
    void setOption(int key, long long value) { check(pp_set_option(h_, key, value), "setOption"); }
    void setOption(int key, double value) { check(pp_set_option_double(h_, key, value), "setOption"); }
    void setStream(void *stream) { check(pp_set_stream(h_, stream), "setStream"); }
    pp_handle handle() const { return h_; }

  private:
    static void check(int rc, const char *where)
    {
        if (rc != PP_SUCCESS) throw PittPackException(rc, where);
    }
    pp_handle h_ = nullptr;
};

// make_Poisson(argc, argv, nx, ny, nz) (src/pencilDcmp.cpp:6432-6440): the reference picks
// its CPU or GPU backend at compile time; here it is the MI355X solver.  argc / argv are not
// read (the reference passes them to MPI).
inline std::unique_ptr<PoissonHIP> make_Poisson(int argc, char **argv, int nx, int ny, int nz)
{
    (void)argc;
    (void)argv;
    return std::unique_ptr<PoissonHIP>(new PoissonHIP(nx, ny, nz));
}

#endif /* PITTPACK_HIP_HPP */
