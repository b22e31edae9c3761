/* make_h5.c - writes the HDF5 fixtures of tests/test_formats.py with the
 * HDF5 library (1.10, from /opt/conda in the build container; run by
 * make_h5.sh).  Every array is a closed formula, restated by the test:
 *   row_ptr[i] = 3 i                       i < 51          (50 rows, 3 each)
 *   col_idx[e] = (7 e + 3) % 50            e < 150
 *   x[i] = 0.25 i - 3, y[i] = 100 - 0.5 i, z[i] = (i % 13) * 1.5, region_id[i] = i % 5   i < 200
 *   big[i] = 1000003 i mod 2^31 (int32), 10,000 entries (chunked and deflated)
 *   f64be[i] = i / 8 (big-endian doubles), i16[i] = i - 20, i64[i] = i * 2^33
 * Layouts: earliest (superblock 0, symbol-table groups, version-1 object
 * headers) and latest (superblock 3, link messages, version-2 object
 * headers); contiguous, compact and chunked (deflate + shuffle) storage; a
 * root group with 40 datasets (several symbol-table nodes) and, in the
 * latest format, one with 20 links (dense link storage).                    */
#include <hdf5.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static void put(hid_t loc, const char* name, hid_t ftype, hid_t mtype, hsize_t n, const void* data, hid_t dcpl) {
    hid_t sp = H5Screate_simple(1, &n, NULL);
    hid_t ds = H5Dcreate2(loc, name, ftype, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT);
    if (ds < 0) { fprintf(stderr, "create %s failed\n", name); exit(1); }
    if (H5Dwrite(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) { fprintf(stderr, "write %s\n", name); exit(1); }
    H5Dclose(ds);
    H5Sclose(sp);
}

static hid_t create(const char* path, int latest) {
    hid_t fapl = H5Pcreate(H5P_FILE_ACCESS);
    if (latest) H5Pset_libver_bounds(fapl, H5F_LIBVER_LATEST, H5F_LIBVER_LATEST);
    else H5Pset_libver_bounds(fapl, H5F_LIBVER_EARLIEST, H5F_LIBVER_LATEST);
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, fapl);
    H5Pclose(fapl);
    if (f < 0) { fprintf(stderr, "cannot create %s\n", path); exit(1); }
    return f;
}

int main(int argc, char** argv) {
    const char* dir = argc > 1 ? argv[1] : ".";
    char path[1024];
    int row_ptr[51], col_idx[150], i16v[40];
    float x[200], y[200], z[200], rid[200];
    double f64[64];
    long long i64[16];
    static int big[10000];
    for (int i = 0; i < 51; ++i) row_ptr[i] = 3 * i;
    for (int e = 0; e < 150; ++e) col_idx[e] = (7 * e + 3) % 50;
    for (int i = 0; i < 200; ++i) {
        x[i] = 0.25f * i - 3.0f;
        y[i] = 100.0f - 0.5f * i;
        z[i] = (float)(i % 13) * 1.5f;
        rid[i] = (float)(i % 5);
    }
    for (int i = 0; i < 64; ++i) f64[i] = i / 8.0;
    for (int i = 0; i < 40; ++i) i16v[i] = i - 20;
    for (int i = 0; i < 16; ++i) i64[i] = (long long)i << 33;
    for (int i = 0; i < 10000; ++i) big[i] = (int)(((long long)1000003 * i) % 2147483648LL);

    for (int latest = 0; latest < 2; ++latest) {
        const char* tag = latest ? "latest" : "earliest";
        snprintf(path, sizeof path, "%s/net_%s.h5", dir, tag);
        hid_t f = create(path, latest);
        put(f, "row_ptr", H5T_STD_I32LE, H5T_NATIVE_INT, 51, row_ptr, H5P_DEFAULT);
        put(f, "col_idx", H5T_STD_I32LE, H5T_NATIVE_INT, 150, col_idx, H5P_DEFAULT);
        H5Fclose(f);

        snprintf(path, sizeof path, "%s/morph_%s.h5", dir, tag);
        f = create(path, latest);
        put(f, "x", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, 200, x, H5P_DEFAULT);
        put(f, "y", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, 200, y, H5P_DEFAULT);
        put(f, "z", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, 200, z, H5P_DEFAULT);
        put(f, "region_id", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, 200, rid, H5P_DEFAULT);
        H5Fclose(f);

        snprintf(path, sizeof path, "%s/types_%s.h5", dir, tag);
        f = create(path, latest);
        hid_t g1 = H5Gcreate2(f, "grp", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        hid_t g2 = H5Gcreate2(g1, "sub", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        put(g2, "f64be", H5T_IEEE_F64BE, H5T_NATIVE_DOUBLE, 64, f64, H5P_DEFAULT);
        put(g1, "i16", H5T_STD_I16LE, H5T_NATIVE_INT, 40, i16v, H5P_DEFAULT);
        put(f, "i64", H5T_STD_I64LE, H5T_NATIVE_LLONG, 16, i64, H5P_DEFAULT);
        hid_t dc = H5Pcreate(H5P_DATASET_CREATE);
        H5Pset_layout(dc, H5D_COMPACT);
        put(f, "compact_f32", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, 200, x, dc);
        H5Pclose(dc);
        dc = H5Pcreate(H5P_DATASET_CREATE);
        hsize_t chunk = 1000;
        H5Pset_chunk(dc, 1, &chunk);
        H5Pset_shuffle(dc);
        H5Pset_deflate(dc, 6);
        put(f, "chunked_i32", H5T_STD_I32LE, H5T_NATIVE_INT, 10000, big, dc);
        H5Pclose(dc);
        dc = H5Pcreate(H5P_DATASET_CREATE);
        H5Pset_chunk(dc, 1, &chunk);
        put(f, "chunked_plain", H5T_STD_I32LE, H5T_NATIVE_INT, 10000, big, dc);
        H5Pclose(dc);
        H5Gclose(g2);
        H5Gclose(g1);
        H5Fclose(f);

        snprintf(path, sizeof path, "%s/many_%s.h5", dir, tag);
        f = create(path, latest);
        const int count = latest ? 20 : 40;
        for (int k = 0; k < count; ++k) {
            char nm[32];
            int v[4] = {k, k + 1, k + 2, k + 3};
            snprintf(nm, sizeof nm, "d%02d", k);
            put(f, nm, H5T_STD_I32LE, H5T_NATIVE_INT, 4, v, H5P_DEFAULT);
        }
        H5Fclose(f);
    }
    printf("fixtures written to %s\n", dir);
    return 0;
}
