// Data loader and program display of the gpuropolis CLI (see io.hpp).
#include "io.hpp"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <queue>

#include "gpuropolis.h"

// Op sets (main.cu:51-86): the heap code of each enum value, in enum order (the
// library's gpr_ref_op_set returns the same table).
static const int kOpsetN[GPR_NUM_OPSETS] = {27, 13, 11, 9, 19};
static const uint8_t kOpset[GPR_NUM_OPSETS][27] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26},
    {GPR_REF_ONE, GPR_REF_CONST, GPR_REF_INV, GPR_REF_A, GPR_REF_B, GPR_REF_PLUS, GPR_REF_MINUS, GPR_REF_RMINUS,
     GPR_REF_CPLUS, GPR_REF_TIMES, GPR_REF_CTIMES, GPR_REF_DIV, GPR_REF_RDIV},
    {GPR_REF_ONE, GPR_REF_CONST, GPR_REF_A, GPR_REF_B, GPR_REF_PLUS, GPR_REF_MINUS, GPR_REF_RMINUS, GPR_REF_CPLUS,
     GPR_REF_TIMES, GPR_REF_CTIMES, GPR_REF_CPOW},
    {GPR_REF_ONE, GPR_REF_CONST, GPR_REF_A, GPR_REF_B, GPR_REF_PLUS, GPR_REF_MINUS, GPR_REF_RMINUS, GPR_REF_CPLUS,
     GPR_REF_CTIMES},
    {GPR_REF_LOGOF2, GPR_REF_HALF, GPR_REF_E, GPR_REF_PI, GPR_REF_CLIGHT, GPR_REF_G, GPR_REF_HBAR, GPR_REF_MU0,
     GPR_REF_EL, GPR_REF_MP, GPR_REF_ME, GPR_REF_E0, GPR_REF_KB, GPR_REF_A, GPR_REF_B, GPR_REF_PLUS, GPR_REF_MINUS,
     GPR_REF_RMINUS, GPR_REF_TIMES}};
static bool is_physical(int o) { return o == GPR_REF_LOGOF2 || (o >= GPR_REF_E && o <= GPR_REF_KB); }

namespace gpr_host {

bool load_data_file(const char *path, float sdscale, bool first_half, bool even_half, std::vector<Datum> &out,
                    std::string &err) {
    FILE *fp = fopen(path, "r");
    if (!fp) {
        err = std::string("*** ERROR: Cannot open file:\t") + path;
        return false;
    }
    out.clear();
    char *line = nullptr;
    size_t cap = 0;
    double x, y, sd;
    bool ok = true;
    while (getline(&line, &cap, fp) != -1) {
        if (line[0] == '#') continue;
        if (sscanf(line, "%lf\t%lf\t%lf\n", &x, &y, &sd) == 3 || sscanf(line, "%le\t%le\t%le\n", &x, &y, &sd) == 3) {
            sd = sdscale * sd;
            if (!((float)sd > 0.0f)) {
                err = std::string("*** ERROR: sd must be positive:\t") + line;
                ok = false;
                break;
            }
            out.push_back(Datum{(float)x, (float)y, (float)sd});
        } else if (strspn(line, " \r\n\t") == strlen(line)) {
            continue;
        } else {
            err = std::string("*** ERROR IN PARSING INPUT\t") + line;
            ok = false;
            break;
        }
    }
    free(line);
    fclose(fp);
    if (!ok) return false;
    if (first_half) out.resize(out.size() / 2);
    if (even_half) {
        std::vector<Datum> kept;
        for (size_t i = 1; i < out.size(); i += 2) kept.push_back(out[i]);
        out.swap(kept);
    }
    return true;
}

static std::string fmt(const char *f, float v) {
    char b[64];
    snprintf(b, sizeof b, f, v);
    return b;
}

// ------------------------------------------------------------------ heap programs
// string_dispatch (main.cu:821-873)
static std::string heap_str_op(int o, const std::string &a, const std::string &b, const std::string &C) {
    switch (o) {
    case GPR_REF_ONE: return "1";
    case GPR_REF_CONST: return C;
    case GPR_REF_INV: return "(1/" + a + ")";
    case GPR_REF_A: return a;
    case GPR_REF_B: return b;
    case GPR_REF_PLUS: return "(" + a + "+" + b + ")";
    case GPR_REF_CPLUS: return "(" + a + "+" + C + ")";
    case GPR_REF_MINUS: return "(" + a + "-" + b + ")";
    case GPR_REF_RMINUS: return "(" + b + "-" + a + ")";
    case GPR_REF_TIMES: return "(" + b + "*" + a + ")";
    case GPR_REF_CTIMES: return "(" + a + "*" + C + ")";
    case GPR_REF_DIV: return "(" + a + "/" + b + ")";
    case GPR_REF_RDIV: return "(" + b + "/" + a + ")";
    case GPR_REF_LOG: return "log(" + a + ")";
    case GPR_REF_SIN: return "sin(" + a + ")";
    case GPR_REF_ASIN: return "asin(" + a + ")";
    case GPR_REF_COS: return "cos(" + a + ")";
    case GPR_REF_ACOS: return "acos(" + a + ")";
    case GPR_REF_TAN: return "tan(" + a + ")";
    case GPR_REF_ATAN: return "atan(" + a + ")";
    case GPR_REF_SQRT: return "sqrt(" + a + ")";
    case GPR_REF_EXP: return "exp(" + a + ")";
    case GPR_REF_POW: return "(" + a + "^" + b + ")";
    case GPR_REF_CPOW: return "(" + a + "^" + C + ")";
    case GPR_REF_RPOW: return "(" + b + "^" + a + ")";
    case GPR_REF_CRPOW: return "(" + C + "^" + a + ")";
    case GPR_REF_ABS: return "abs(" + a + ")";
    case GPR_REF_E: return "E";
    case GPR_REF_PI: return "PI";
    case GPR_REF_CLIGHT: return "CLIGHT";
    case GPR_REF_G: return "G";
    case GPR_REF_HBAR: return "HBAR";
    case GPR_REF_MU0: return "MU0";
    case GPR_REF_EL: return "EL";
    case GPR_REF_MP: return "MP";
    case GPR_REF_ME: return "ME";
    case GPR_REF_E0: return "E0";
    case GPR_REF_KB: return "KB";
    case GPR_REF_LOGOF2: return "log(2.0)";
    case GPR_REF_HALF: return "(" + a + "/2.0)";
    default: return "NAN";
    }
}
// latex_dispatch (main.cu:901-953)
static std::string heap_tex_op(int o, const std::string &a, const std::string &b, const std::string &C) {
    switch (o) {
    case GPR_REF_ONE: return "1";
    case GPR_REF_CONST: return C;
    case GPR_REF_INV: return "1/" + a;
    case GPR_REF_A: return a;
    case GPR_REF_B: return b;
    case GPR_REF_PLUS: return a + "+" + b;
    case GPR_REF_CPLUS: return a + "+" + C;
    case GPR_REF_MINUS: return a + "-" + b;
    case GPR_REF_RMINUS: return b + "-" + a;
    case GPR_REF_TIMES: return b + " " + a;
    case GPR_REF_CTIMES: return a + " \\cdot " + C;
    case GPR_REF_DIV: return "\\frac{" + a + "}{" + b + "}";
    case GPR_REF_RDIV: return "\\frac{" + b + "}{" + a + "}";
    case GPR_REF_LOG: return "\\log(" + a + ")";
    case GPR_REF_SIN: return "sin(" + a + ")";
    case GPR_REF_ASIN: return "asin(" + a + ")";
    case GPR_REF_COS: return "cos(" + a + ")";
    case GPR_REF_ACOS: return "acos(" + a + ")";
    case GPR_REF_TAN: return "tan(" + a + ")";
    case GPR_REF_ATAN: return "atan(" + a + ")";
    case GPR_REF_SQRT: return "\\sqrt{" + a + "}";
    case GPR_REF_EXP: return "e^{" + a + "}";
    case GPR_REF_POW: return a + "^{" + b + "}";
    case GPR_REF_CPOW: return a + "^{" + C + "}";
    case GPR_REF_RPOW: return b + "^{" + a + "}";
    case GPR_REF_CRPOW: return C + "^{" + a + "}";
    case GPR_REF_ABS: return "\\mid " + a + "\\mid ";
    case GPR_REF_E: return "e";
    case GPR_REF_PI: return "\\pi";
    case GPR_REF_CLIGHT: return "c";
    case GPR_REF_G: return "G";
    case GPR_REF_HBAR: return "\\bar{h}";
    case GPR_REF_MU0: return "\\mu_0";
    case GPR_REF_EL: return "e_l";
    case GPR_REF_MP: return "m_p";
    case GPR_REF_ME: return "m_e";
    case GPR_REF_E0: return "\\epsilon_0";
    case GPR_REF_KB: return "k_B";
    case GPR_REF_LOGOF2: return "\\log 2)";               // unbalanced as in main.cu:947
    case GPR_REF_HALF: return "(" + a + "/2.0)";
    default: return "NAN";
    }
}

// bottom-up string fold over the 31 slots, x below slots 16..31 (main.cu:878-898)
template <class F>
static std::string heap_fold_str(const uint8_t *P, const float *C, const char *cf, F op) {
    std::vector<std::string> buf(2 * GPR_REF_SLOTS + 2);
    for (int i = GPR_REF_SLOTS; i >= 1; i--) {
        const int l = 2 * i, r = 2 * i + 1;
        const std::string a = l > GPR_REF_SLOTS ? "x" : buf[l];
        const std::string b = r > GPR_REF_SLOTS ? "x" : buf[r];
        buf[i] = op(P[i - 1], a, b, fmt(cf, C[i - 1]));
    }
    return buf[1];
}
std::string heap_string(const uint8_t *P, const float *C, const char *cf) { return heap_fold_str(P, C, cf, heap_str_op); }
std::string heap_latex(const uint8_t *P, const float *C, const char *cf) { return heap_fold_str(P, C, cf, heap_tex_op); }

float heap_length(const uint8_t *P) {  // dispatch_length fold, main.cu:477-526
    float buf[GPR_REF_SLOTS + 2];
    for (int i = GPR_REF_SLOTS; i >= 1; i--) {
        const int l = 2 * i;
        const float a = l > GPR_REF_SLOTS ? 1.0f : buf[l], b = l > GPR_REF_SLOTS ? 1.0f : buf[l + 1];
        float v;
        switch (P[i - 1]) {
        case GPR_REF_ONE: case GPR_REF_CONST: v = 1.0f; break;
        case GPR_REF_A: v = a; break;
        case GPR_REF_B: v = b; break;
        case GPR_REF_CPLUS: case GPR_REF_CTIMES: case GPR_REF_CPOW: case GPR_REF_CRPOW: v = 1.0f + a + 1.0f; break;
        case GPR_REF_PLUS: case GPR_REF_MINUS: case GPR_REF_TIMES: case GPR_REF_DIV: case GPR_REF_RDIV:
        case GPR_REF_POW: v = 1.0f + a + b; break;
        case GPR_REF_RMINUS: case GPR_REF_RPOW: v = 1.0f + b + a; break;
        case GPR_REF_INV: case GPR_REF_SQRT: case GPR_REF_LOG: case GPR_REF_SIN: case GPR_REF_ASIN: case GPR_REF_COS:
        case GPR_REF_ACOS: case GPR_REF_TAN: case GPR_REF_ATAN: case GPR_REF_EXP: case GPR_REF_ABS:
        case GPR_REF_HALF: v = 1.0f + a; break;
        default: v = is_physical(P[i - 1]) ? 1.0f : NAN; break;   // PHYSICAL_CONSTANT_LENGTH_PRIOR
        }
        buf[i] = v;
    }
    return buf[1];
}

static float fmaxr(float a, float b) { return a > b ? a : b; }  // max macro, main.cu:184-187
float heap_degree(const uint8_t *P, const float *C) {  // dispatch_degree fold, main.cu:615-654
    float buf[GPR_REF_SLOTS + 2];
    for (int i = GPR_REF_SLOTS; i >= 1; i--) {
        const int l = 2 * i;
        const float a = l > GPR_REF_SLOTS ? 1.0f : buf[l], b = l > GPR_REF_SLOTS ? 1.0f : buf[l + 1];
        const float c = C[i - 1];
        float v;
        switch (P[i - 1]) {
        case GPR_REF_ONE: case GPR_REF_CONST: v = 0.0f; break;
        case GPR_REF_A: v = a; break;
        case GPR_REF_B: v = b; break;
        case GPR_REF_PLUS: case GPR_REF_MINUS: case GPR_REF_RMINUS: v = fmaxr(a, b); break;
        case GPR_REF_CPLUS: case GPR_REF_CTIMES: v = a; break;
        case GPR_REF_TIMES: v = a + b; break;
        case GPR_REF_CPOW: v = a * c; break;
        case GPR_REF_POW: v = b == 0.0f ? a * c : INFINITY; break;
        case GPR_REF_DIV: v = b == 0.0f ? a : INFINITY; break;
        case GPR_REF_HALF: v = a; break;
        default: v = is_physical(P[i - 1]) ? 0.0f : INFINITY; break;
        }
        buf[i] = v;
    }
    return buf[1];
}

// PROGRAM_CODE of each op set (main.cu:55, 61, 67, 74, 85), indexed by enum value
const char *opset_program_code(int set) {
    switch (set) {
    case GPR_OPS_SIMPLIFIED: return "1CIab+-_#*@/|";
    case GPR_OPS_POLYNOMIAL: return "1CAB+-_#*@^";
    case GPR_OPS_LINEAR: return "1CAB+-_#@";
    case GPR_OPS_PHYSICAL: return "B2epcghm:o.~Kab+-_*";
    default: return GPR_REF_PROGRAM_CODE;
    }
}
std::string heap_code(const uint8_t *P, int set) {
    if (set < 0 || set >= GPR_NUM_OPSETS) set = GPR_OPS_DEFAULT;
    const char *code = opset_program_code(set);
    std::string s;
    for (int i = 0; i < GPR_REF_SLOTS; i++) {
        char ch = '?';                                  // a code outside the op set
        for (int k = 0; k < kOpsetN[set]; k++)
            if (kOpset[set][k] == P[i]) ch = code[k];
        s += ch;
    }
    return s;
}

// ------------------------------------------------------------------ postfix programs
static std::string tape_str_op(int o, const std::string &a, const std::string &b, const std::string &C, bool tex) {
    switch (o) {
    case GPR_OP_ADD: return tex ? a + "+" + b : "(" + a + "+" + b + ")";
    case GPR_OP_SUB: return tex ? a + "-" + b : "(" + a + "-" + b + ")";
    case GPR_OP_MUL: return tex ? a + " " + b : "(" + a + "*" + b + ")";
    case GPR_OP_DIV: return tex ? "\\frac{" + a + "}{" + b + "}" : "(" + a + "/" + b + ")";
    case GPR_OP_POW: return tex ? a + "^{" + b + "}" : "(" + a + "^" + b + ")";
    case GPR_OP_RSUB: return tex ? b + "-" + a : "(" + b + "-" + a + ")";
    case GPR_OP_RDIV: return tex ? "\\frac{" + b + "}{" + a + "}" : "(" + b + "/" + a + ")";
    case GPR_OP_RPOW: return tex ? b + "^{" + a + "}" : "(" + b + "^" + a + ")";
    case GPR_OP_NEG: return tex ? "-" + a : "(-" + a + ")";
    case GPR_OP_EXP: return tex ? "e^{" + a + "}" : "exp(" + a + ")";
    case GPR_OP_LOG: return tex ? "\\log(" + a + ")" : "log(" + a + ")";
    case GPR_OP_INV: return tex ? "1/" + a : "(1/" + a + ")";
    case GPR_OP_SQRT: return tex ? "\\sqrt{" + a + "}" : "sqrt(" + a + ")";
    case GPR_OP_SIN: return "sin(" + a + ")";
    case GPR_OP_ASIN: return "asin(" + a + ")";
    case GPR_OP_COS: return "cos(" + a + ")";
    case GPR_OP_ACOS: return "acos(" + a + ")";
    case GPR_OP_TAN: return "tan(" + a + ")";
    case GPR_OP_ATAN: return "atan(" + a + ")";
    case GPR_OP_ABS: return tex ? "\\mid " + a + "\\mid " : "abs(" + a + ")";
    case GPR_OP_CADD: return tex ? a + "+" + C : "(" + a + "+" + C + ")";
    case GPR_OP_CMUL: return tex ? a + " \\cdot " + C : "(" + a + "*" + C + ")";
    case GPR_OP_CPOW: return tex ? a + "^{" + C + "}" : "(" + a + "^" + C + ")";
    case GPR_OP_CRPOW: return tex ? C + "^{" + a + "}" : "(" + C + "^" + a + ")";
    default: return "NAN";
    }
}

static std::string tape_fold(const uint8_t *ops, const float *cs, int len, const char *cf, bool tex) {
    std::vector<std::string> st;
    for (int t = 0; t < len; t++) {
        const int o = ops[t];
        if (o == GPR_OP_X) st.push_back("x");
        else if (o == GPR_OP_CONST) st.push_back(fmt(cf, cs[t]));
        else if (o == GPR_OP_ONE) st.push_back("1");
        else if (o >= GPR_OP_ADD && o <= GPR_OP_RPOW) {
            if (st.size() < 2) return "NAN";
            std::string b = st.back(); st.pop_back();
            std::string a = st.back(); st.pop_back();
            st.push_back(tape_str_op(o, a, b, "", tex));
        } else {
            if (st.empty()) return "NAN";
            std::string a = st.back(); st.pop_back();
            st.push_back(tape_str_op(o, a, "", fmt(cf, cs[t]), tex));
        }
    }
    return st.size() == 1 ? st[0] : "NAN";
}
std::string tape_string(const uint8_t *ops, const float *cs, int len, const char *cf) { return tape_fold(ops, cs, len, cf, false); }
std::string tape_latex(const uint8_t *ops, const float *cs, int len, const char *cf) { return tape_fold(ops, cs, len, cf, true); }

float tape_length(const uint8_t *ops, int len) {
    float l = 0.0f;
    for (int t = 0; t < len; t++) l += ops[t] >= GPR_OP_CADD ? 2.0f : 1.0f;
    return l;
}

// polynomial degree of a postfix program: x 1, constants 0, + - max, * sum, / and ^ only by
// constant divisors / exponents, constant-operand forms as in dispatch_degree
float tape_degree(const uint8_t *ops, const float *cs, int len) {
    std::vector<float> st;
    std::vector<int> isc;
    std::vector<float> cval;
    for (int t = 0; t < len; t++) {
        const int o = ops[t];
        if (o == GPR_OP_X) { st.push_back(1.0f); isc.push_back(0); cval.push_back(0.0f); continue; }
        if (o == GPR_OP_CONST || o == GPR_OP_ONE) {
            st.push_back(0.0f); isc.push_back(1); cval.push_back(o == GPR_OP_ONE ? 1.0f : cs[t]);
            continue;
        }
        if (o >= GPR_OP_ADD && o <= GPR_OP_RPOW) {
            if (st.size() < 2) return INFINITY;
            const float b = st.back(), a = st[st.size() - 2];
            const int cb = isc.back();
            const float vb = cval.back();
            st.pop_back(); isc.pop_back(); cval.pop_back();
            float v;
            switch (o) {
            case GPR_OP_ADD: case GPR_OP_SUB: case GPR_OP_RSUB: v = fmaxr(a, b); break;
            case GPR_OP_MUL: v = a + b; break;
            case GPR_OP_DIV: v = b == 0.0f ? a : INFINITY; break;
            case GPR_OP_POW: v = (cb && vb >= 0.0f && vb == floorf(vb)) ? a * vb : INFINITY; break;
            default: v = INFINITY; break;
            }
            st.back() = v; isc.back() = 0; cval.back() = 0.0f;
            continue;
        }
        if (st.empty()) return INFINITY;
        float &a = st.back();
        switch (o) {
        case GPR_OP_NEG: case GPR_OP_CADD: case GPR_OP_CMUL: break;
        case GPR_OP_CPOW: a = (cs[t] >= 0.0f && cs[t] == floorf(cs[t])) ? a * cs[t] : INFINITY; break;
        default: a = INFINITY; break;
        }
        isc.back() = 0;
    }
    return st.size() == 1 ? st[0] : INFINITY;
}

std::string tape_code(const uint8_t *ops, int len) {
    // the reference's symbols where an op exists there (PROGRAM_CODE, main.cu:55), x and ~ (neg)
    static const char code[GPR_NUM_OPS + 1] = "xC1+-*/^_|V~ELIRSsCcTtA#@pP";
    std::string s;
    for (int t = 0; t < len; t++) s += ops[t] < GPR_NUM_OPS ? code[ops[t]] : '?';
    return s;
}

// ------------------------------------------------------------------ tops.txt (run-all.sh:45)
namespace {

struct TopLine {
    int cls;         // 0 key does not parse, 1 NaN, 2 number
    long double v;   // the key when cls == 2
    std::string s;   // the line without its newline
};

TopLine top_key(std::string s) {
    const char *p = s.c_str(), *e = p + s.size();
    auto blank = [](char c) { return c == ' ' || c == '\t'; };
    for (int f = 0; f < 2 && p < e; f++) {   // skip fields 1 and 2 (blanks, then non-blanks)
        while (p < e && blank(*p)) p++;
        while (p < e && !blank(*p)) p++;
    }
    TopLine t{0, 0.0L, std::move(s)};
    char *end = nullptr;
    const long double v = strtold(p, &end);
    if (end != p) {
        t.cls = std::isnan(v) ? 1 : 2;
        t.v = v;
    }
    return t;
}

// strict "a sorts before b" of `sort -g -k3` with its whole-line last resort
bool top_less(const TopLine &a, const TopLine &b) {
    if (a.cls != b.cls) return a.cls < b.cls;
    if (a.cls == 2 && a.v != b.v) return a.v < b.v;
    // sort orders NaNs by their bit pattern: the quiet NaN strtold returns for "nan"
    // before the one it returns for "-nan"
    if (a.cls == 1 && std::signbit(a.v) != std::signbit(b.v)) return !std::signbit(a.v);
    return a.s < b.s;   // std::string compares bytes as unsigned char, as memcmp does
}

struct TopGreater {
    bool operator()(const TopLine &a, const TopLine &b) const { return top_less(b, a); }
};

}  // namespace

bool write_tops(const char *samples, const char *out, long k, std::string &err) {
    std::ifstream in(samples, std::ios::binary);
    if (!in) {
        err = std::string("cannot open ") + samples;
        return false;
    }
    // min-heap of the k largest lines seen so far
    std::priority_queue<TopLine, std::vector<TopLine>, TopGreater> heap;
    std::string line;
    while (k > 0 && std::getline(in, line)) {
        TopLine t = top_key(std::move(line));
        if ((long)heap.size() < k) heap.push(std::move(t));
        else if (top_less(heap.top(), t)) {
            heap.pop();
            heap.push(std::move(t));
        }
        line.clear();
    }
    std::vector<TopLine> v;
    v.reserve(heap.size());
    while (!heap.empty()) {
        v.push_back(heap.top());
        heap.pop();
    }
    FILE *fp = fopen(out, "w");
    if (!fp) {
        err = std::string("cannot write ") + out;
        return false;
    }
    bool ok = true;
    for (const TopLine &t : v) ok = ok && fwrite(t.s.data(), 1, t.s.size(), fp) == t.s.size() && fputc('\n', fp) != EOF;
    return fclose(fp) == 0 && ok;
}

}  // namespace gpr_host
