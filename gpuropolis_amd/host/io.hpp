// Host-side I/O of the gpuropolis CLI: the data loader, program display strings and
// the samples.txt / log.txt writers.  Formats follow the reference's run loop
// (main.cu:312-363 loader, 818-976 display, 1087-1211 samples.txt).
#pragma once
#include <stdint.h>
#include <stdio.h>

#include <string>
#include <vector>

namespace gpr_host {

struct Datum {
    float x, y, sd;
};

// load_data_file (main.cu:319-363): '#' comments, "x y sd" triples (any whitespace,
// %lf then %le), blank lines skipped, anything else is an error; sd *= sdscale and must
// stay > 0; optional first-half / even-half trimming.  Returns false with a message.
bool load_data_file(const char *path, float sdscale, bool first_half, bool even_half, std::vector<Datum> &out,
                    std::string &err);

// Heap-31 program display (reference semantics): P/C are one chain's 31 slots.
std::string heap_string(const uint8_t *P, const float *C, const char *const_format);   // displaystring
std::string heap_latex(const uint8_t *P, const float *C, const char *const_format);    // display_latex
float heap_length(const uint8_t *P);                                                   // dispatch_length fold
float heap_degree(const uint8_t *P, const float *C);                                   // polynomial_degree
std::string heap_code(const uint8_t *P, int opset = 0);                               // PROGRAM_CODE chars (op set)
const char *opset_program_code(int opset);                                            // main.cu:55-85

// Postfix program display (config semantics).
std::string tape_string(const uint8_t *ops, const float *cs, int len, const char *const_format);
std::string tape_latex(const uint8_t *ops, const float *cs, int len, const char *const_format);
float tape_length(const uint8_t *ops, int len);
float tape_degree(const uint8_t *ops, const float *cs, int len);
std::string tape_code(const uint8_t *ops, int len);

// tops.txt: the k lines of `samples` with the largest third field, in increasing order,
// as run-all.sh:45 makes it with `sort -g -k3 samples.txt | tail -n 10000` (C locale).
// Keys follow sort's general-numeric rule: fields start at a blank-to-non-blank
// transition, the key is strtold of field 3 with its leading blanks, lines whose key does
// not parse come first, then NaN, then numbers in order; equal keys fall back to a
// bytewise comparison of the whole line.  Only k lines are held in memory.
bool write_tops(const char *samples, const char *out, long k, std::string &err);

}  // namespace gpr_host
