// The .PUR settings file (source/purification_settings.f90:91-218), parsed on
// the host: the drop-in's hosts keep reading the same input files.
//
// Grammar, as the reference reads it:
// * one line per record, truncated to 80 characters (character(len=80) line);
// * tokens are separated by blanks and tabs (source/readline.f90); a token
//   whose first ten characters read as a real under F10.5 is a number (its
//   value is the whole token), 'true' / 'false' are logicals, anything else is
//   a string (checktype, source/readl2.f90:52-92);
// * the first string starting with '!' marks a comment line;
// * "key = value": the second string is '=', the third the value (strings) or
//   the first number (numeric keys, truncated by int() - so
//   chempot_increment 0.1 becomes 0, as in the reference); a key that is not
//   recognised is ignored; an unknown value of a string key is the
//   reference's error stop (here an error status with its message);
// * a line whose second string is not '=' sets the logical named by its first
//   string (dev, check, iterative_inversion).
// Keys the reference does not read (the chemical-potential guess, the metric's
// cycle count) keep their defaults here too.
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ptbgpu.h"

namespace {

// Does the token read as a real under Fortran F10.5 (its first ten
// characters: sign, digits with an optional point, an optional exponent with
// letter E/D/Q or a bare sign, or NaN / Inf)?
bool reads_as_real(const std::string &tok) {
    std::string f = tok.substr(0, 10);
    size_t i = 0, n = f.size();
    if (n == 0) return false;
    std::string low;
    for (char c : f) low += (char)std::tolower((unsigned char)c);
    size_t s = (low[0] == '+' || low[0] == '-') ? 1 : 0;
    const std::string body = low.substr(s);
    if (body == "nan" || body == "inf" || body == "infinity" || body.rfind("nan(", 0) == 0)
        return true;
    i = s;
    size_t digits = 0;
    while (i < n && std::isdigit((unsigned char)low[i])) ++i, ++digits;
    if (i < n && low[i] == '.') {
        ++i;
        while (i < n && std::isdigit((unsigned char)low[i])) ++i, ++digits;
    }
    if (digits == 0) return false;
    if (i == n) return true;
    if (low[i] == 'e' || low[i] == 'd' || low[i] == 'q') ++i;
    else if (low[i] != '+' && low[i] != '-') return false;
    if (i < n && (low[i] == '+' || low[i] == '-')) ++i;
    size_t ed = 0;
    while (i < n && std::isdigit((unsigned char)low[i])) ++i, ++ed;
    return i == n && ed > 0;
}

double real_value(const std::string &tok) {
    std::string t = tok;
    for (char &c : t)
        if (c == 'd' || c == 'D' || c == 'q' || c == 'Q') c = 'e';
    return std::strtod(t.c_str(), nullptr);
}

void set_error(char *msg, int64_t len, const char *text) {
    if (msg && len > 0) {
        std::strncpy(msg, text, (size_t)len - 1);
        msg[len - 1] = '\0';
    }
}

}  // namespace

extern "C" {

int ptbgpu_pur_defaults(ptbgpu_pur_file *f) {
    if (!f) return PTBGPU_E_ARG;
    std::memset(f, 0, sizeof(*f));
    f->mode = 1;                 // full
    f->submatrix_columns = 2;    // atoms_columns
    f->submatrix_mode = 2;       // submatrix_sygv
    f->prlvl = 1;
    f->dev = 1;
    f->check = 0;
    f->pur.type = 4;             // sign_diagonalization
    f->pur.metric = 2;           // inv_sqrt
    f->pur.metric_iterative = 0;
    f->pur.cycles = 40;
    f->pur.metric_cycles = 50;
    f->pur.chempot_cycles = 60;
    f->pur.chempot_guess = -1.5;
    f->pur.chempot_increment = 0.1;
    f->pur.nel = 0;
    return PTBGPU_OK;
}

int ptbgpu_pur_parse(const char *text, int64_t len, ptbgpu_pur_file *f, char *errmsg,
                     int64_t errlen) {
    if (!f || (!text && len > 0) || len < 0) return PTBGPU_E_ARG;
    ptbgpu_pur_defaults(f);
    set_error(errmsg, errlen, "");
    int64_t pos = 0;
    while (pos < len) {
        int64_t end = pos;
        while (end < len && text[end] != '\n') ++end;
        std::string line(text + pos, text + end);
        pos = end + 1;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.size() > 80) line.resize(80);
        std::vector<std::string> strs;
        std::vector<double> nums;
        size_t i = 0;
        while (i < line.size()) {
            while (i < line.size() && (line[i] == ' ' || line[i] == '\t')) ++i;
            size_t j = i;
            while (j < line.size() && line[j] != ' ' && line[j] != '\t') ++j;
            if (j > i) {
                const std::string tok = line.substr(i, j - i);
                if (reads_as_real(tok)) nums.push_back(real_value(tok));
                else if (tok != "true" && tok != "false") strs.push_back(tok);
            }
            i = j;
        }
        const std::string a1 = strs.size() > 0 ? strs[0] : "";
        if (!a1.empty() && a1[0] == '!') continue;
        const std::string a2 = strs.size() > 1 ? strs[1] : "";
        if (a2 == "=") {
            const std::string a3 = strs.size() > 2 ? strs[2] : "";
            auto pick = [&](const char *const *names, const int64_t *codes, int count,
                            int64_t &dst, const char *err) -> bool {
                for (int k = 0; k < count; ++k)
                    if (a3 == names[k]) {
                        dst = codes[k];
                        return true;
                    }
                set_error(errmsg, errlen, err);
                return false;
            };
            static const char *kInvType = "Error: .PUR contains invalid type definition";
            if (a1 == "mode") {
                static const char *n[] = {"full", "submatrix"};
                static const int64_t c[] = {1, 2};
                if (!pick(n, c, 2, f->mode, kInvType)) return PTBGPU_E_ARG;
            } else if (a1 == "submatrix_columns") {
                static const char *n[] = {"single_columns", "atoms_columns", "full_matrix",
                                          "kmeans_r_heuristic"};
                static const int64_t c[] = {1, 2, 3, 4};
                if (!pick(n, c, 4, f->submatrix_columns, kInvType)) return PTBGPU_E_ARG;
            } else if (a1 == "submatrix_mode") {
                static const char *n[] = {"submatrix_sign", "submatrix_sygv"};
                static const int64_t c[] = {1, 2};
                if (!pick(n, c, 2, f->submatrix_mode, kInvType)) return PTBGPU_E_ARG;
            } else if (a1 == "type") {
                static const char *n[] = {"mcweeny", "pade", "newton", "diag"};
                static const int64_t c[] = {1, 2, 3, 4};
                if (!pick(n, c, 4, f->pur.type, kInvType)) return PTBGPU_E_ARG;
            } else if (a1 == "metric") {
                static const char *n[] = {"inverse", "inverse_sqrt", "sqrt"};
                static const int64_t c[] = {1, 2, 3};
                if (!pick(n, c, 3, f->pur.metric, "Error: .PUR contains invalid metric definition."))
                    return PTBGPU_E_ARG;
            } else if (!nums.empty()) {
                // int() truncates toward zero
                const double v = std::trunc(nums[0]);
                if (a1 == "cycles") f->pur.cycles = (int64_t)v;
                else if (a1 == "chempot_cycles") f->pur.chempot_cycles = (int64_t)v;
                else if (a1 == "chempot_increment") f->pur.chempot_increment = v;
                else if (a1 == "pr") f->prlvl = (int64_t)v;
            }
        } else if (a1 == "dev") {
            f->dev = 1;
        } else if (a1 == "check") {
            f->check = 1;
        } else if (a1 == "iterative_inversion") {
            f->pur.metric_iterative = 1;
        }
    }
    return PTBGPU_OK;
}

}  // extern "C"
