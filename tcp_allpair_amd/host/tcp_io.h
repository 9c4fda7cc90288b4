/*
 * tcp_io.h -- host-side readers and writers of the reference's text files, so
 * that a standalone C++ driver and the unchanged Fortran driver (main.f90 +
 * result.f90 over the GPU drop-in) read the same inputs and write the same
 * bytes.  C ABI (also loaded by the Python tests).
 *
 *   tcp.prm      main.f90:221-291   positional list-directed records
 *   restart.xyz  main.f90:383-438   2 header records, then name x y z vx vy vz fx fy fz id
 *   ener##.dat   main.f90:37-41, result.f90:319-324  FORMAT(5(ES15.8, 1X))
 *   RDF###.dat   result.f90:243-252                  FORMAT(6(ES14.6, 1X))
 *   VDF###.dat   result.f90:256-260                  FORMAT(4(ES14.6, 1X))
 *   REST##.xyz   result.f90:74-131                   FORMAT(A3, 9(1X, ES14.6), 1X, I6)
 *
 * Numbers are printed with the semantics of Fortran's ESw.d edit descriptor
 * (exponent letter dropped for |exponent| > 99) and list-directed records as
 * the flang runtime writes them (a leading blank, items separated by one blank).
 */
#ifndef TCP_IO_H
#define TCP_IO_H

#ifdef __cplusplus
extern "C" {
#endif

/* tcp.prm (main.f90:221-291).  Times in fs as written in the file. */
typedef struct tcp_prm {
    double tmax, trest, tdump, tsamp, dt; /* fs */
    double box[3];
    double Te, Ti;
    int NTel, NTion;
    double rc, rs;
    double xm[2], q[2];
    char thermo[8];                       /* 'NOTEMP' 'BEREND' 'ISOKIN' 'STOCH ' */
    double tau, alpha;                    /* Berendsen tau, Langevin friction */
    char rest[8], trace[8];
    double dx[3];                         /* RDF bin widths ee, ei, ii */
    int K[3];
} tcp_prm;

/* 0 ok, -1 cannot open, -2 malformed (err receives a message of at most 256 bytes) */
int tcp_io_read_prm(const char* path, tcp_prm* p, char* err);
/* restart.xyz of n = NTel + NTion particles: AoS x, v, f [3n] and ids [n] */
int tcp_io_read_restart(const char* path, int n, double* x, double* v, double* f, int* id,
                        char* err);

/* Fortran ESw.d of v into out (w + 1 bytes); returns w. */
int tcp_io_format_es(double v, int w, int d, char* out);

/* File writers; the directory is the current one.  Return 0 or -1. */
/* ener##.dat: the header record (main.f90:37-41); append one row of SAMPLING. */
int tcp_io_ener_open(const char* path);
int tcp_io_ener_row(const char* path, double t_fs, double Tel, double Tion, double Epot, double Etot);
/* RDF###.dat and VDF###.dat of DUMP for counts rdf[3*501], vdf[2*501] gathered over Ndump steps */
int tcp_io_write_dump(int nfreq, int ntel, int ntion, int ndump, const double* box, const double* dx,
                      double dv_e, double dv_i, const double* rdf, const double* vdf);
/* REST##.xyz of Restart */
int tcp_io_write_rest(int nfreq, int ntel, int ntion, const double* x, const double* v,
                      const double* f, const int* id);

#ifdef __cplusplus
}
#endif
#endif
