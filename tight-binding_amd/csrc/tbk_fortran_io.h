// tbk_fortran_io.h -- Fortran-style input of text files: the formatted (F, D,
// I, L, A, X edit descriptors) and list-directed reads with which the
// reference's program reads its input files, with their field semantics:
//   * each READ statement starts at the next record; short records are
//     padded with blanks;
//   * Fw.d without a decimal point carries d implied decimals; D exponents
//     and exponents written as a bare sign (1.0+05) are accepted;
//   * a unit opened with BLANK='ZERO' (the SKIN file, static.f:682) reads the
//     non-leading blanks of a numeric field as zeros, so "12   " under I5 is
//     12000; otherwise blanks are ignored;
//   * list-directed input takes values separated by blanks or commas over as
//     many records as it needs and discards the rest of the last one; a
//     repeat count r*v, a null value (two commas, r*) and '/' are honoured.
#pragma once

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <string>

namespace tbk {

class FortranReader {
public:
    explicit FortranReader(FILE *f, bool blank_zero = false) : f_(f), bz_(blank_zero) {}

    // Start a READ statement: the next record (false at end of file).
    bool next_record()
    {
        line_.clear();
        pos_ = 0;
        rep_ = 0;
        slash_ = false;
        int ch;
        bool any = false;
        while ((ch = std::fgetc(f_)) != EOF) {
            any = true;
            if (ch == '\n') break;
            if (ch != '\r') line_.push_back((char)ch);
        }
        if (any) ++records_;
        return any;
    }
    long records() const { return records_; }
    const std::string &record() const { return line_; }

    // ---- formatted input on the current record --------------------------
    std::string field(int w)
    {
        std::string s;
        for (int i = 0; i < w; ++i, ++pos_) s.push_back(pos_ < line_.size() ? line_[pos_] : ' ');
        return s;
    }
    void skip(int w) { pos_ += (size_t)w; }
    // Aw: the characters with trailing blanks removed
    std::string chars(int w)
    {
        std::string s = field(w);
        while (!s.empty() && (s.back() == ' ' || s.back() == '\t')) s.pop_back();
        return s;
    }
    // Fw.d / Dw.d / Ew.d
    bool real(int w, int d, double *out) { return to_real(numeric(field(w)), d, out); }
    // Iw
    bool integer(int w, int *out)
    {
        const std::string s = numeric(field(w));
        if (s.empty()) { *out = 0; return true; }
        char *end = nullptr;
        const long v = std::strtol(s.c_str(), &end, 10);
        *out = (int)v;
        return end && *end == '\0';
    }
    // Lw: optional blanks and a period, then T or F (anything after ignored)
    bool logical(int w, int *out) { return to_logical(field(w), out); }
    // a list of n reals under a format of per_rec fields Fw.d per record
    bool real_list(int n, int per_rec, int w, int d, double *out)
    {
        for (int i = 0; i < n; ++i) {
            if (i % per_rec == 0 && !next_record()) return false;
            if (!real(w, d, out + i)) return false;
        }
        return true;
    }

    // ---- list-directed input (call next_record() first) ------------------
    // Each returns false on a read error or end of file; a null value or a
    // value after '/' leaves *out unchanged and returns true.
    bool list_real(double *out)
    {
        std::string t;
        bool null = false;
        if (!token(&t, &null)) return false;
        return null || to_real(t, 0, out);
    }
    bool list_int(int *out)
    {
        std::string t;
        bool null = false;
        if (!token(&t, &null)) return false;
        if (null) return true;
        char *end = nullptr;
        const long v = std::strtol(t.c_str(), &end, 10);
        if (t.empty() || !end || *end != '\0') return false;
        *out = (int)v;
        return true;
    }
    bool list_logical(int *out)
    {
        std::string t;
        bool null = false;
        if (!token(&t, &null)) return false;
        return null || to_logical(t, out);
    }

private:
    // the characters of a numeric field after the blank rule of the unit
    std::string numeric(const std::string &f) const
    {
        std::string s;
        bool lead = true;
        for (char c : f) {
            if (c == ' ' || c == '\t') {
                if (!lead && bz_) s.push_back('0');
                continue;
            }
            lead = false;
            s.push_back(c);
        }
        return s;
    }
    static bool to_real(std::string s, int d, double *out)
    {
        for (char &c : s)
            if (c == 'd' || c == 'D' || c == 'q' || c == 'Q') c = 'E';
        if (s.empty()) { *out = 0.0; return true; }
        const bool has_point = s.find('.') != std::string::npos;
        size_t e = s.find_first_of("Ee");
        if (e == std::string::npos) {   // exponent written as a bare sign: 1.0+05
            for (size_t i = 1; i < s.size(); ++i)
                if ((s[i] == '+' || s[i] == '-') && (std::isdigit((unsigned char)s[i - 1]) || s[i - 1] == '.')) {
                    s.insert(i, "E");
                    e = i;
                    break;
                }
        }
        if (!has_point && d > 0) {
            std::string mant = e == std::string::npos ? s : s.substr(0, e);
            long ex = e == std::string::npos ? 0 : std::strtol(s.c_str() + e + 1, nullptr, 10);
            s = mant + "E" + std::to_string(ex - d);
        }
        char *end = nullptr;
        *out = std::strtod(s.c_str(), &end);
        return end && *end == '\0';
    }
    static bool to_logical(const std::string &f, int *out)
    {
        size_t i = 0;
        while (i < f.size() && (f[i] == ' ' || f[i] == '\t')) ++i;
        if (i < f.size() && f[i] == '.') ++i;
        if (i >= f.size()) return false;
        const char c = (char)std::toupper((unsigned char)f[i]);
        if (c != 'T' && c != 'F') return false;
        *out = c == 'T';
        return true;
    }
    static bool sep(char c) { return c == ' ' || c == '\t' || c == ','; }
    // The next value of a list-directed read: *null set for a null value
    // (empty between commas, r*, or after '/').
    bool token(std::string *t, bool *null)
    {
        *null = false;
        t->clear();
        if (slash_) { *null = true; return true; }
        if (rep_ > 0) {
            --rep_;
            *t = rep_value_;
            *null = rep_null_;
            return true;
        }
        for (;;) {
            while (pos_ < line_.size() && (line_[pos_] == ' ' || line_[pos_] == '\t')) ++pos_;
            if (pos_ >= line_.size()) {                 // the value list continues on the next record
                if (!next_record_keep()) return false;
                continue;
            }
            break;
        }
        if (line_[pos_] == '/') { slash_ = true; *null = true; return true; }
        if (line_[pos_] == ',') {                        // empty value between separators
            ++pos_;
            *null = true;
            return true;
        }
        size_t b = pos_;
        if (line_[pos_] == '\'' || line_[pos_] == '"') {    // a quoted string
            const char q = line_[pos_++];
            while (pos_ < line_.size() && line_[pos_] != q) t->push_back(line_[pos_++]);
            if (pos_ < line_.size()) ++pos_;
        } else {
            while (pos_ < line_.size() && !sep(line_[pos_]) && line_[pos_] != '/') ++pos_;
            *t = line_.substr(b, pos_ - b);
        }
        // consume one separator: blanks and at most one comma
        while (pos_ < line_.size() && (line_[pos_] == ' ' || line_[pos_] == '\t')) ++pos_;
        if (pos_ < line_.size() && line_[pos_] == ',') ++pos_;
        const size_t star = t->find('*');
        if (star != std::string::npos && star > 0 &&
            t->find_first_not_of("0123456789") == star) {   // r*value or r*
            const int r = std::atoi(t->substr(0, star).c_str());
            rep_value_ = t->substr(star + 1);
            rep_null_ = rep_value_.empty();
            rep_ = r - 1;
            *t = rep_value_;
            *null = rep_null_;
        }
        return true;
    }
    bool next_record_keep()
    {
        const int rep = rep_;
        const bool sl = slash_;
        const bool ok = next_record();
        rep_ = rep;
        slash_ = sl;
        return ok;
    }

    FILE *f_;
    bool bz_;
    std::string line_;
    size_t pos_ = 0;
    long records_ = 0;
    int rep_ = 0;
    bool rep_null_ = false, slash_ = false;
    std::string rep_value_;
};

}  // namespace tbk
