// pp_cxx_check.cpp -- main.cpp's call order (src/main.cpp:73-118) written against the C++ class
// of include/pittpack_hip.hpp, built by g++ into libpp_cxx_check.so so a test can run the C++
// binding itself.  Returns the code of the first failing call (PittPackException), 0 on success.
#include <cstring>

#include "pittpack_hip.hpp"

extern "C" int pp_cxx_main_case(int nx, int ny, int nz, const char *bc, const double *X, const double *rhs,
                                double *u, double *err)
{
    try {
        auto M = make_Poisson(0, nullptr, nx, ny, nz);
        char mybc[6];
        std::memcpy(mybc, bc, 6);
        M->assignBoundary(mybc);
        double box[6];
        std::memcpy(box, X, sizeof(box));
        M->setBox(box);
        M->setCoords(2);
        if (rhs) M->assignRhs(rhs);
        else M->initializeTrigonometric();
        M->pittPack();
        *err = M->getError();
        M->getSolution(u);
        return PP_SUCCESS;
    } catch (const PittPackException &e) {
        return e.code();
    } catch (...) {
        return -1;
    }
}
