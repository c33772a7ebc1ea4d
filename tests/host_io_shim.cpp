// C entry points over the CLI's host I/O (gpuropolis_amd/host/io.cpp) so the CPU tests
// can call the loader and the display functions through ctypes.  Test infrastructure:
// compiled by tests/test_cli_cpu.py together with io.cpp.
#include <string.h>

#include "../gpuropolis_amd/host/io.hpp"

using namespace gpr_host;

static int put(const std::string &s, char *buf, int cap) {
    if ((int)s.size() + 1 > cap) return -1;
    memcpy(buf, s.c_str(), s.size() + 1);
    return (int)s.size();
}

extern "C" {

// returns the number of points written to xyz (3 floats each), or -1 with err filled
int shim_load(const char *path, float sdscale, int first_half, int even_half, float *xyz, int maxn, char *err,
              int errcap) {
    std::vector<Datum> d;
    std::string e;
    if (!load_data_file(path, sdscale, first_half, even_half, d, e)) {
        put(e, err, errcap);
        return -1;
    }
    if ((int)d.size() > maxn) return -2;
    for (size_t i = 0; i < d.size(); i++) {
        xyz[3 * i] = d[i].x;
        xyz[3 * i + 1] = d[i].y;
        xyz[3 * i + 2] = d[i].sd;
    }
    return (int)d.size();
}

// which: 0 string, 1 latex, 2 code
int shim_heap(const uint8_t *P, const float *C, int which, const char *cf, char *buf, int cap) {
    if (which == 0) return put(heap_string(P, C, cf), buf, cap);
    if (which == 1) return put(heap_latex(P, C, cf), buf, cap);
    return put(heap_code(P), buf, cap);
}
int shim_heap_code(const uint8_t *P, int opset, char *buf, int cap) { return put(heap_code(P, opset), buf, cap); }
float shim_heap_length(const uint8_t *P) { return heap_length(P); }
float shim_heap_degree(const uint8_t *P, const float *C) { return heap_degree(P, C); }

int shim_tops(const char *in, const char *out, long k, char *err, int errcap) {
    std::string e;
    if (!write_tops(in, out, k, e)) {
        put(e, err, errcap);
        return -1;
    }
    return 0;
}

int shim_tape(const uint8_t *ops, const float *cs, int len, int which, const char *cf, char *buf, int cap) {
    if (which == 0) return put(tape_string(ops, cs, len, cf), buf, cap);
    if (which == 1) return put(tape_latex(ops, cs, len, cf), buf, cap);
    return put(tape_code(ops, len), buf, cap);
}
float shim_tape_length(const uint8_t *ops, int len) { return tape_length(ops, len); }
float shim_tape_degree(const uint8_t *ops, const float *cs, int len) { return tape_degree(ops, cs, len); }
}
