// pp_output.hip -- solution output, the counterpart of PencilDcmp::IO (src/pencilDcmp.cpp:1827-1832)
// and of the writer it calls, Phdf5::writeMultiBlockCellCenter + xdmfMultiBlockCellCenter
// (src/phdf5.cpp:1299-1808, 1809-2237).
//
// Default format (PP_IO_HDF5): "<folder>/Pxdmf3d<app>.h5" with four FP64 datasets in the
// reference's shapes -- X, Y, Z: node coordinates on the (nz+1) x (ny+1) x (nx+1) cell corners,
// Q: the cell-centred field, nz x ny x nx -- each with a trailing block dimension of 1 (the
// reference's 4th dimension counts chunks; one device holds one block) and chunked as one
// block, plus the XDMF 2.0 index "<folder>/Pxdmf3d<app>.xmf" (a spatial collection holding one
// 3DSMesh grid whose X_Y_Z geometry and cell attribute Q are hyperslabs of those datasets).
// HDF5 is loaded at run time (dlopen of the serial libhdf5, 1.10 API), so the library has no
// link dependency on it; without it PP_IO_HDF5 returns PP_NO_HDF5.
//
// PP_IO_RAW: the field as raw little-endian FP64 "<folder>/Pxdmf3d<app>.bin" (z, y, x order)
// with an XDMF index describing the same mesh as a 3DCoRectMesh (origin and spacing).
#include <dlfcn.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <utility>
#include <string>
#include <vector>

#include "pittpack_hip.h"
#include "pp_handle.h"

namespace {

// The subset of the HDF5 1.10 C API the writer uses (hid_t is int64_t from 1.10 on).
typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;
constexpr hid_t kH5Default = 0;  // H5P_DEFAULT, H5S_ALL
constexpr unsigned kH5AccTrunc = 0x0002u;

struct H5Api {
    bool ok = false;
    herr_t (*open)();
    herr_t (*libversion)(unsigned *, unsigned *, unsigned *);
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
    herr_t (*sclose)(hid_t);
    hid_t (*pcreate)(hid_t);
    herr_t (*pset_chunk)(hid_t, int, const hsize_t *);
    herr_t (*pclose)(hid_t);
    hid_t (*dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    herr_t (*dclose)(hid_t);
    herr_t (*eset_auto2)(hid_t, void *, void *);
    hid_t *cls_dataset_create;  // H5P_CLS_DATASET_CREATE_ID_g
    hid_t *native_double;       // H5T_NATIVE_DOUBLE_g
};

template <class F>
bool bind(void *lib, const char *name, F &f)
{
    f = reinterpret_cast<F>(dlsym(lib, name));
    return f != nullptr;
}

const H5Api &h5()
{
    static H5Api api;
    static std::once_flag once;
    std::call_once(once, [] {
        std::vector<std::string> names;
        if (const char *env = std::getenv("PP_HDF5_LIB")) names.push_back(env);
        for (const char *n : {"libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103", "/opt/conda/lib/libhdf5.so",
                              "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so"})
            names.push_back(n);
        void *lib = nullptr;
        for (const auto &n : names)
            if ((lib = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
        if (!lib) return;
        H5Api a;
        bool ok = bind(lib, "H5open", a.open) && bind(lib, "H5get_libversion", a.libversion) &&
                  bind(lib, "H5Fcreate", a.fcreate) && bind(lib, "H5Fclose", a.fclose) &&
                  bind(lib, "H5Screate_simple", a.screate_simple) && bind(lib, "H5Sclose", a.sclose) &&
                  bind(lib, "H5Pcreate", a.pcreate) && bind(lib, "H5Pset_chunk", a.pset_chunk) &&
                  bind(lib, "H5Pclose", a.pclose) && bind(lib, "H5Dcreate2", a.dcreate2) &&
                  bind(lib, "H5Dwrite", a.dwrite) && bind(lib, "H5Dclose", a.dclose);
        a.cls_dataset_create = reinterpret_cast<hid_t *>(dlsym(lib, "H5P_CLS_DATASET_CREATE_ID_g"));
        a.native_double = reinterpret_cast<hid_t *>(dlsym(lib, "H5T_NATIVE_DOUBLE_g"));
        ok = ok && a.cls_dataset_create && a.native_double;
        unsigned maj = 0, min = 0, rel = 0;
        // the hid_t width above is the 1.10+ one; the class ids are valid after H5open
        ok = ok && a.libversion(&maj, &min, &rel) >= 0 && (maj > 1 || (maj == 1 && min >= 10)) && a.open() >= 0;
        if (!ok) return;
        // failures are returned as PP_IO_FAIL; the library's own error-stack printing stays off
        if (bind(lib, "H5Eset_auto2", a.eset_auto2)) a.eset_auto2(kH5Default, nullptr, nullptr);
        a.ok = true;
        api = a;
    });
    return api;
}

// one 4-D FP64 dataset {d0, d1, d2, 1}, chunked as one block, written whole
bool h5_dataset(const H5Api &a, hid_t file, const char *name, hsize_t d0, hsize_t d1, hsize_t d2, const double *v)
{
    const hsize_t dims[4] = {d0, d1, d2, 1};
    const hid_t space = a.screate_simple(4, dims, nullptr);
    if (space < 0) return false;
    const hid_t dcpl = a.pcreate(*a.cls_dataset_create);
    bool ok = dcpl >= 0 && a.pset_chunk(dcpl, 4, dims) >= 0;
    hid_t ds = -1;
    if (ok) ds = a.dcreate2(file, name, *a.native_double, space, kH5Default, dcpl, kH5Default);
    ok = ok && ds >= 0 && a.dwrite(ds, *a.native_double, kH5Default, kH5Default, kH5Default, v) >= 0;
    if (ds >= 0) ok = a.dclose(ds) >= 0 && ok;
    if (dcpl >= 0) a.pclose(dcpl);
    a.sclose(space);
    return ok;
}

std::string path_in(const char *folder, const std::string &file)
{
    const std::string dir(folder);
    return dir.empty() ? file : dir + "/" + file;
}

int write_hdf5(const char *folder, int app, int nx, int ny, int nz, const double *X, const double *u)
{
    const H5Api &a = h5();
    if (!a.ok) return PP_NO_HDF5;
    const std::string stem = "Pxdmf3d" + std::to_string(app);
    const size_t L = (size_t)nx + 1, M = (size_t)ny + 1, N = (size_t)nz + 1;
    std::vector<double> c;
    try {
        c.resize(L * M * N);
    } catch (const std::bad_alloc &) {
        return PP_ALLOCATION_FAIL;
    }
    const hid_t file = a.fcreate(path_in(folder, stem + ".h5").c_str(), kH5AccTrunc, kH5Default, kH5Default);
    if (file < 0) return PP_IO_FAIL;
    // node coordinates: origin + spacing * index, spacing (b - a) / n (getXcoord / getYcoord /
    // getZcoord with aligndir 0, src/phdf5.cpp:2238-2420)
    const double h[3] = {(X[1] - X[0]) / nx, (X[3] - X[2]) / ny, (X[5] - X[4]) / nz};
    bool ok = true;
    const char *names[3] = {"/X", "/Y", "/Z"};
    for (int d = 0; d < 3 && ok; d++) {
        size_t o = 0;
        for (size_t k = 0; k < N; k++)
            for (size_t j = 0; j < M; j++)
                for (size_t i = 0; i < L; i++) {
                    const size_t idx = d == 0 ? i : d == 1 ? j : k;
                    c[o++] = X[2 * d] + h[d] * (double)idx;
                }
        ok = h5_dataset(a, file, names[d], N, M, L, c.data());
    }
    ok = ok && h5_dataset(a, file, "/Q", (hsize_t)nz, (hsize_t)ny, (hsize_t)nx, u);
    ok = a.fclose(file) >= 0 && ok;
    if (!ok) return PP_IO_FAIL;

    FILE *f = std::fopen(path_in(folder, stem + ".xmf").c_str(), "w");
    if (!f) return PP_IO_FAIL;
    const std::string h5name = stem + ".h5";
    // a hyperslab of block 0 of a {d0 d1 d2 1} dataset
    auto slab = [&](const char *name, size_t d0, size_t d1, size_t d2) {
        return std::fprintf(f,
                            "        <DataItem ItemType=\"HyperSlab\" Dimensions=\"%zu %zu %zu 1\" Type=\"HyperSlab\">\n"
                            "          <DataItem Dimensions=\"3 4\" Format=\"XML\">0 0 0 0  1 1 1 1  %zu %zu %zu 1</DataItem>\n"
                            "          <DataItem Name=\"%s\" Dimensions=\"%zu %zu %zu 1\" NumberType=\"Float\" "
                            "Precision=\"8\" Format=\"HDF\">%s:/%s</DataItem>\n"
                            "        </DataItem>\n",
                            d0, d1, d2, d0, d1, d2, name, d0, d1, d2, h5name.c_str(), name) >= 0;
    };
    bool w = std::fprintf(f,
                          "<?xml version=\"1.0\" ?>\n"
                          "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n"
                          "<Xdmf Version=\"2.0\">\n"
                          "<Domain>\n"
                          "<Grid Name=\"AMR\" GridType=\"Collection\" CollectionType=\"Spatial\">\n"
                          "  <Grid Name=\"mesh0\" GridType=\"Uniform\">\n"
                          "    <Topology TopologyType=\"3DSMesh\" NumberOfElements=\"%zu %zu %zu\"/>\n"
                          "    <Geometry GeometryType=\"X_Y_Z\">\n",
                          N, M, L) >= 0;
    for (const char *n : {"X", "Y", "Z"}) w = w && slab(n, N, M, L);
    w = w && std::fprintf(f, "    </Geometry>\n    <Attribute Name=\"Q\" AttributeType=\"Scalar\" Center=\"Cell\">\n") >= 0;
    w = w && slab("Q", (size_t)nz, (size_t)ny, (size_t)nx);
    w = w && std::fprintf(f, "    </Attribute>\n  </Grid>\n</Grid>\n</Domain>\n</Xdmf>\n") >= 0;
    if (std::fclose(f) != 0 || !w) return PP_IO_FAIL;
    return PP_SUCCESS;
}

int write_raw(const char *folder, int app, int nx, int ny, int nz, const double *X, const double *u)
{
    const size_t n = (size_t)nx * ny * nz;
    const std::string stem = "Pxdmf3d" + std::to_string(app);
    const unsigned probe = 1;
    unsigned char first;
    std::memcpy(&first, &probe, 1);
    FILE *f = std::fopen(path_in(folder, stem + ".bin").c_str(), "wb");
    if (!f) return PP_IO_FAIL;
    bool ok = true;
    if (first == 1) {
        ok = std::fwrite(u, sizeof(double), n, f) == n;
    } else {  // big-endian host: the file is little-endian
        for (size_t i = 0; i < n && ok; i++) {
            unsigned char b[8];
            std::memcpy(b, &u[i], 8);
            for (int k = 0; k < 4; k++) std::swap(b[k], b[7 - k]);
            ok = std::fwrite(b, 1, 8, f) == 8;
        }
    }
    if (std::fclose(f) != 0 || !ok) return PP_IO_FAIL;
    const double dx = (X[1] - X[0]) / nx, dy = (X[3] - X[2]) / ny, dz = (X[5] - X[4]) / nz;
    f = std::fopen(path_in(folder, stem + ".xmf").c_str(), "w");
    if (!f) return PP_IO_FAIL;
    int w = std::fprintf(f,
                         "<?xml version=\"1.0\" ?>\n"
                         "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n"
                         "<Xdmf Version=\"2.0\">\n"
                         "  <Domain>\n"
                         "    <Grid Name=\"P\" GridType=\"Uniform\">\n"
                         "      <Topology TopologyType=\"3DCoRectMesh\" Dimensions=\"%d %d %d\"/>\n"
                         "      <Geometry GeometryType=\"ORIGIN_DXDYDZ\">\n"
                         "        <DataItem Name=\"Origin\" Dimensions=\"3\" NumberType=\"Float\" Precision=\"8\" "
                         "Format=\"XML\">%.17g %.17g %.17g</DataItem>\n"
                         "        <DataItem Name=\"Spacing\" Dimensions=\"3\" NumberType=\"Float\" Precision=\"8\" "
                         "Format=\"XML\">%.17g %.17g %.17g</DataItem>\n"
                         "      </Geometry>\n"
                         "      <Attribute Name=\"Q\" AttributeType=\"Scalar\" Center=\"Cell\">\n"
                         "        <DataItem Dimensions=\"%d %d %d\" NumberType=\"Float\" Precision=\"8\" "
                         "Format=\"Binary\" Endian=\"Little\">%s.bin</DataItem>\n"
                         "      </Attribute>\n"
                         "    </Grid>\n"
                         "  </Domain>\n"
                         "</Xdmf>\n",
                         nz + 1, ny + 1, nx + 1, X[4], X[2], X[0], dz, dy, dx, nz, ny, nx, stem.c_str());
    if (std::fclose(f) != 0 || w < 0) return PP_IO_FAIL;
    return PP_SUCCESS;
}

}  // namespace

extern "C" {

int pp_write_solution(const char *folder, int app, int nx, int ny, int nz, const double *X, const double *u, int format)
{
    if (!folder || !X || !u || app < 0 || nx < 1 || ny < 1 || nz < 1) return PP_INPUT_ARGS;
    if (!(X[1] > X[0] && X[3] > X[2] && X[5] > X[4])) return PP_INPUT_ARGS;
    if (format == PP_IO_HDF5) return write_hdf5(folder, app, nx, ny, nz, X, u);
    if (format == PP_IO_RAW) return write_raw(folder, app, nx, ny, nz, X, u);
    return PP_INPUT_ARGS;
}

int pp_hdf5_available(void) { return h5().ok ? 1 : 0; }

int pp_io(pp_handle h, const char *folder, int app)
{
    if (!h) return PP_INVALID_HANDLE;
    if (!folder || app < 0) return PP_INPUT_ARGS;
    if (!h->box_set || !h->coords_set) return PP_CALL_ORDER;
    const size_t n = (size_t)h->g.nx * h->g.ny * h->g.nz;
    std::vector<double> u;
    try {
        u.resize(n);
    } catch (const std::bad_alloc &) {
        return PP_ALLOCATION_FAIL;
    }
    if (hipMemcpyAsync(u.data(), h->d_field, n * sizeof(double), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return PP_HIP_ERROR;
    return pp_write_solution(folder, app, h->g.nx, h->g.ny, h->g.nz, h->g.X, u.data(), h->io_format);
}

}  // extern "C"
