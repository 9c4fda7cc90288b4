// Readers and writers of the reference's text files (declarations and the
// file/format map in tcp_io.h).
#include "tcp_io.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace {

constexpr double kPi = 3.1415926535897931;  // data.f90:23
constexpr int kNdist = 500;                 // data.f90:16

void set_err(char* err, const std::string& m)
{
    if (err) {
        std::strncpy(err, m.c_str(), 255);
        err[255] = '\0';
    }
}

// List-directed input: values separated by blanks or commas, quoted strings,
// a READ that needs more items than a record holds continues on the next one.
class ListReader {
public:
    explicit ListReader(std::istream& in) : in_(in) {}

    // the items of one READ statement
    bool read(int n, std::vector<std::string>& out)
    {
        out.clear();
        std::string line;
        while ((int)out.size() < n) {
            if (!std::getline(in_, line)) return false;
            split(line, out, n);
        }
        return true;
    }

private:
    static void split(const std::string& s, std::vector<std::string>& out, int n)
    {
        size_t i = 0;
        while (i < s.size() && (int)out.size() < n) {
            while (i < s.size() && (std::isspace((unsigned char)s[i]) || s[i] == ',')) ++i;
            if (i >= s.size()) break;
            if (s[i] == '\'' || s[i] == '"') {
                const char q = s[i++];
                std::string t;
                while (i < s.size() && s[i] != q) t += s[i++];
                ++i;
                out.push_back(t);
            } else {
                std::string t;
                while (i < s.size() && !std::isspace((unsigned char)s[i]) && s[i] != ',') t += s[i++];
                out.push_back(t);
            }
        }
    }
    std::istream& in_;
};

bool to_double(const std::string& t, double* v)
{
    std::string s = t;
    for (char& c : s)
        if (c == 'd' || c == 'D') c = 'E';
    char* end = nullptr;
    *v = std::strtod(s.c_str(), &end);
    return end && *end == '\0' && !s.empty();
}

bool to_int(const std::string& t, int* v)
{
    char* end = nullptr;
    const long x = std::strtol(t.c_str(), &end, 10);
    *v = (int)x;
    return end && *end == '\0' && !t.empty();
}

void copy_chars(char* dst, const std::string& s, size_t cap)
{
    std::memset(dst, 0, cap);
    std::strncpy(dst, s.c_str(), cap - 1);
}

std::string es(double v, int w, int d)
{
    char buf[64];
    tcp_io_format_es(v, w, d, buf);
    return buf;
}

// Fortran I2.2 / I3.3 of a file number (asterisks when it does not fit)
std::string zpad(int v, int w)
{
    char buf[32];
    std::snprintf(buf, sizeof buf, "%0*d", w, v);
    std::string s(buf);
    if ((int)s.size() > w || v < 0) return std::string(w, '*');
    return s;
}

}  // namespace

extern "C" {

int tcp_io_format_es(double v, int w, int d, char* out)
{
    std::string s;
    if (std::isnan(v)) {
        s = "NaN";
    } else if (std::isinf(v)) {
        s = v < 0 ? "-Infinity" : "Infinity";
    } else {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.*E", d, v);
        s = buf;
        const size_t e = s.find('E');
        const int ex = std::atoi(s.c_str() + e + 1);
        char eb[16];
        if (ex > 99 || ex < -99)
            std::snprintf(eb, sizeof eb, "%c%03d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
        else
            std::snprintf(eb, sizeof eb, "E%c%02d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
        s = s.substr(0, e) + eb;
    }
    if ((int)s.size() > w) s = std::string(w, '*');
    else s = std::string(w - s.size(), ' ') + s;
    std::memcpy(out, s.c_str(), (size_t)w + 1);
    return w;
}

int tcp_io_read_prm(const char* path, tcp_prm* p, char* err)
{
    std::ifstream in(path);
    if (!in) {
        set_err(err, std::string("cannot open ") + path);
        return -1;
    }
    std::memset(p, 0, sizeof *p);
    ListReader r(in);
    std::vector<std::string> t;
    bool ok = true;
    auto dbl = [&](int n, double* v) {
        if (!(ok = ok && r.read(n, t))) return;
        for (int i = 0; i < n && ok; ++i) ok = to_double(t[i], &v[i]);
    };
    auto num = [&](int n, int* v) {
        if (!(ok = ok && r.read(n, t))) return;
        for (int i = 0; i < n && ok; ++i) ok = to_int(t[i], &v[i]);
    };
    auto skip = [&]() { ok = ok && r.read(1, t); };
    double tt[5];
    skip();                                   // title
    skip(); dbl(5, tt);                       // tmax trest tdump tsamp dt
    p->tmax = tt[0]; p->trest = tt[1]; p->tdump = tt[2]; p->tsamp = tt[3]; p->dt = tt[4];
    skip(); dbl(3, p->box);
    skip(); dbl(1, &p->Te); dbl(1, &p->Ti);
    skip(); num(1, &p->NTel); num(1, &p->NTion);
    skip(); dbl(1, &p->rc);
    skip(); dbl(1, &p->rs);
    skip(); dbl(1, &p->xm[0]); dbl(1, &p->xm[1]);
    skip(); dbl(1, &p->q[0]); dbl(1, &p->q[1]);
    skip();
    if (ok && (ok = r.read(1, t))) copy_chars(p->thermo, (t[0] + "      ").substr(0, 6), sizeof p->thermo);
    double ta[2];
    dbl(2, ta);
    p->tau = ta[0];
    p->alpha = ta[1];
    skip();
    if (ok && (ok = r.read(2, t))) {
        copy_chars(p->rest, t[0], sizeof p->rest);
        copy_chars(p->trace, t[1], sizeof p->trace);
    }
    skip(); dbl(3, p->dx);
    skip(); num(3, p->K);
    if (!ok) {
        set_err(err, std::string("malformed ") + path);
        return -2;
    }
    return 0;
}

int tcp_io_read_restart(const char* path, int n, double* x, double* v, double* f, int* id, char* err)
{
    std::ifstream in(path);
    if (!in) {
        set_err(err, std::string("cannot open ") + path);
        return -1;
    }
    ListReader r(in);
    std::vector<std::string> t;
    if (!r.read(1, t) || !r.read(1, t)) {
        set_err(err, std::string("malformed header in ") + path);
        return -2;
    }
    for (int i = 0; i < n; ++i) {
        double a[9];
        bool ok = r.read(11, t);
        for (int k = 0; k < 9 && ok; ++k) ok = to_double(t[1 + k], &a[k]);
        ok = ok && to_int(t[10], &id[i]);
        if (!ok) {
            set_err(err, "malformed particle record " + std::to_string(i + 1) + " in " + path);
            return -2;
        }
        for (int k = 0; k < 3; ++k) {
            x[3 * (size_t)i + k] = a[k];
            v[3 * (size_t)i + k] = a[3 + k];
            f[3 * (size_t)i + k] = a[6 + k];
        }
    }
    return 0;
}

int tcp_io_ener_open(const char* path)
{
    FILE* fp = std::fopen(path, "w");
    if (!fp) return -1;
    std::fputs("# time(fs),       T_el(eV),      T_ion(eV), Potential Energy,  Total energy, "
               "Electron VAF, Ion VAF,  mutual VAF\n", fp);
    return std::fclose(fp) ? -1 : 0;
}

int tcp_io_ener_row(const char* path, double t_fs, double Tel, double Tion, double Epot, double Etot)
{
    FILE* fp = std::fopen(path, "a");
    if (!fp) return -1;
    const double v[5] = {t_fs, Tel, Tion, Epot, Etot};
    // nX only moves the position: no blank after the last item of a record
    std::string s;
    for (int i = 0; i < 5; ++i) s += (i ? " " : "") + es(v[i], 15, 8);
    s += "\n";
    std::fputs(s.c_str(), fp);
    return std::fclose(fp) ? -1 : 0;
}

int tcp_io_write_dump(int nfreq, int ntel, int ntion, int ndump, const double* box, const double* dx,
                      double dv_e, double dv_i, const double* rdf, const double* vdf)
{
    const double Ntau = (double)ndump;
    const double V = box[0] * box[1] * box[2];
    const int ntpt = ntel + ntion;
    const double cnt[3] = {(double)ntel, (double)ntpt, (double)ntion};
    double dV[3];
    for (int t = 0; t < 3; ++t) {
        const double n2 = cnt[t] * cnt[t];
        const double d3 = dx[t] * dx[t] * dx[t];
        dV[t] = n2 * Ntau * d3 * 4.0 * kPi / 3.0 / V;
    }
    std::string name = "RDF" + zpad(nfreq, 3) + ".dat";
    FILE* fp = std::fopen(name.c_str(), "w");
    if (!fp) return -1;
    std::fputs("#  x_ee,      RDF_ee,      x_ei    RDF_ei,       x_ii     RDF_ii\n", fp);
    for (int i = 1; i <= kNdist; ++i) {
        const double shell = (double)(i * i * i - (i - 1) * (i - 1) * (i - 1));
        std::string s;
        for (int t = 0; t < 3; ++t) {
            s += (t ? " " : "") + es((double)i * dx[t], 14, 6);
            s += " " + es(rdf[t * (kNdist + 1) + i - 1] / dV[t] / shell, 14, 6);
        }
        s += "\n";
        std::fputs(s.c_str(), fp);
    }
    if (std::fclose(fp)) return -1;
    name = "VDF" + zpad(nfreq, 3) + ".dat";
    fp = std::fopen(name.c_str(), "w");
    if (!fp) return -1;
    std::fputs("#  T_e        VDF_e            T_i      VDF_i\n", fp);
    for (int i = 1; i <= kNdist; ++i) {
        std::string s = es(dv_e * (double)i, 14, 6) + " " + es(vdf[i - 1] / Ntau, 14, 6) + " " +
                        es(dv_i * (double)i, 14, 6) + " " +
                        es(vdf[(kNdist + 1) + i - 1] / Ntau, 14, 6) + "\n";
        std::fputs(s.c_str(), fp);
    }
    return std::fclose(fp) ? -1 : 0;
}

int tcp_io_write_rest(int nfreq, int ntel, int ntion, const double* x, const double* v,
                      const double* f, const int* id)
{
    const std::string name = "REST" + zpad(nfreq, 2) + ".xyz";
    FILE* fp = std::fopen(name.c_str(), "w");
    if (!fp) return -1;
    // list-directed records: a leading blank, items separated by one blank
    std::fprintf(fp, " %d\n", ntel + ntion);
    std::fprintf(fp, " frame =  %d energy = 0\n", nfreq);
    for (int i = 0; i < ntel + ntion; ++i) {
        std::string s = i < ntel ? "Cu " : " C ";
        const double* src[3] = {x, v, f};
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) s += " " + es(src[a][3 * (size_t)i + k], 14, 6);
        char ib[32];
        std::snprintf(ib, sizeof ib, " %6d\n", id[i]);
        s += ib;
        std::fputs(s.c_str(), fp);
    }
    return std::fclose(fp) ? -1 : 0;
}

}  // extern "C"
