// Field files of the reference, on the host:
//
//  print_su3_soa_ASCII / read_su3_soa_ASCII        OpenAcc/io.c:35-207
//  print_su3_soa_ildg_binary / read_..._binary      OpenAcc/io.c:257-547 (LIME records, ILDG 1.0)
//    with the site and direction order of           OpenAcc/io.c:553-614
//  print_vec3_soa / read_vec3_soa                   DbgTools/dbgtools.c:1193-1236
//
// The reference's functions work on host arrays (global_su3_soa: the nine
// colour planes of each of the 8 link arrays, no K plane; vec3_soa) that its
// wrappers fill from or copy to the device; these do the same on host
// memory, and a device address is staged through a host copy.  The text
// layouts (printf formats), the record order, the big-endian byte order, the
// 8-byte record padding and the determinant test of the ASCII reader are the
// reference's, so files pass between the two programs.  ostp_save_conf and
// ostp_read_conf add the reference's save/read wrappers for the library's
// own (device or registered) su3_soa arrays.
#include "ostp_runtime.h"

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

extern "C" {
// Include/setting_file_parser.c:38: the input file echoed into ILDG files
char input_file_str[300 * 500];
}

namespace {

using namespace ostp;
using cd = std::complex<double>;

// extents for host-only work: the ones given to ostp_set_lattice_host since
// the last ostp_init, else the device context's
const Lattice* io_lattice(const char* nm) {
    if (ctx().host_lat_set) return &ctx().host_lat;
    if (ctx().initialized) return &ctx().lat;
    fail("%s: GEOMETRY MUST BE INITIALIZED (ostp_init or ostp_set_lattice_host)", nm);
    return nullptr;
}

// physical extents nx, ny, nz, nt: nd[xyztmap[k]] (geometry.c:120-131)
void phys_extents(const Lattice& L, int n[4]) {
    for (int k = 0; k < 4; k++) n[k] = L.nd[L.map[k]];
}

size_t snum_h(const Lattice& L, const int d[4]) {
    return (size_t)((d[0] + L.nd[0] * (d[1] + L.nd[1] * (d[2] + (size_t)L.nd[2] * d[3]))) >> 1);
}

// a host view of `bytes` at p: p itself for host memory (registered or not),
// a staged copy for device memory
struct HostView {
    char* p = nullptr;
    void* dev = nullptr;
    std::vector<char> buf;
    bool ok = false;
    HostView(const void* ptr, size_t bytes, bool copy_in, const char* nm) {
        if (!ptr) {
            fail("%s: null pointer", nm);
            return;
        }
        hipPointerAttribute_t a;
        bool device = false;
        if (ctx().initialized) {
            if (hipPointerGetAttributes(&a, ptr) == hipSuccess)
                device = a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
            (void)hipGetLastError();
        }
        if (!device) {
            p = (char*)ptr;
            ok = true;
            return;
        }
        dev = const_cast<void*>(ptr);
        buf.resize(bytes);
        p = buf.data();
        // on the library stream, after the kernels queued there
        ok = !copy_in || ostp_memcpy_dtoh(p, ptr, bytes) == 0;
        (void)nm;
    }
    bool flush(size_t bytes, const char* nm) {
        (void)nm;
        return !dev || ostp_memcpy_htod(dev, p, bytes) == 0;
    }
};

// element (r, c) of link array l at half-site i in a plane-major array with
// link arrays `stride` bytes apart
inline cd& at(char* base, size_t stride, size_t sizeh, int l, int r, int c, size_t i) {
    return ((cd*)(base + (size_t)l * stride))[(size_t)(3 * r + c) * sizeh + i];
}

void rebuild3(cd m[3][3]) {
    m[2][0] = std::conj(m[0][1] * m[1][2] - m[0][2] * m[1][1]);
    m[2][1] = std::conj(m[0][2] * m[1][0] - m[0][0] * m[1][2]);
    m[2][2] = std::conj(m[0][0] * m[1][1] - m[0][1] * m[1][0]);
}

double det_re(const cd m[3][3]) {
    const cd d = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                 m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    return d.real();
}

bool little_endian() {
    const uint16_t one = 1;
    return *(const char*)&one == 1;
}
template <typename T>
T big_endian(T v) {
    if (!little_endian()) return v;
    char* p = (char*)&v;
    for (size_t k = 0; k < sizeof(T) / 2; k++) std::swap(p[k], p[sizeof(T) - 1 - k]);
    return v;
}

// LIME record header (io.c:249-255): 144 bytes, big endian
struct LimeHeader {
    uint32_t magic_no;
    uint16_t version;
    uint16_t mbme_flag;
    uint64_t data_length;
    char type[128];
};
static_assert(sizeof(LimeHeader) == 144, "LIME header layout");

bool write_record(FILE* fp, const char* type, const char* data, uint64_t len) {
    const uint64_t pad = len % 8 == 0 ? 0 : 8 - len % 8;
    LimeHeader h;
    memset(&h, 0, sizeof h);
    h.magic_no = big_endian<uint32_t>(0x456789ab);
    h.version = big_endian<uint16_t>(1);
    h.mbme_flag = 0;
    h.data_length = big_endian<uint64_t>(len + pad);  // the reference records the padded length
    strncpy(h.type, type, sizeof h.type - 1);
    static const char zeros[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    return fwrite(&h, sizeof h, 1, fp) == 1 && (len == 0 || fwrite(data, 1, len, fp) == len) &&
           (pad == 0 || fwrite(zeros, 1, pad, fp) == pad);
}

// the site and direction order of rw_iterate_on_global_sites_lx_xyzt_axis_ordering:
// t, z, y, x (x fastest), then the directions x, y, z, t of each site
template <typename F>
void ildg_sites(const Lattice& L, F&& f) {
    int n[4];
    phys_extents(L, n);
    int d0123[4];
    for (int k = 0; k < 4; k++) d0123[L.map[k]] = k;
    for (int t = 0; t < n[3]; t++)
        for (int z = 0; z < n[2]; z++)
            for (int y = 0; y < n[1]; y++)
                for (int x = 0; x < n[0]; x++) {
                    const int xs[4] = {x, y, z, t};
                    const int parity = (x + y + z + t) % 2;
                    int d[4];
                    for (int k = 0; k < 4; k++) d[k] = xs[d0123[k]];
                    const size_t i = snum_h(L, d);
                    for (int dir = 0; dir < 4; dir++) f(2 * L.map[dir] + parity, i);
                }
}

}  // namespace

extern "C" {

void rw_iterate_on_global_sites_lx_xyzt_axis_ordering(
    void (*single_element_rw)(const int, const int, const int, const void*, const int, FILE*), const void* datastruct,
    FILE* fp, const int scalar_even_mode) {
    const char* nm = "rw_iterate_on_global_sites_lx_xyzt_axis_ordering";
    const Lattice* L = io_lattice(nm);
    if (!L) return;
    if (!single_element_rw) {
        fail("%s: null callback", nm);
        return;
    }
    int n[4];
    phys_extents(*L, n);
    int d0123[4];
    for (int k = 0; k < 4; k++) d0123[L->map[k]] = k;
    const int swap = little_endian() ? 1 : 0;
    const int xlimit = scalar_even_mode ? n[0] / 2 : n[0];
    const int dirlimit = scalar_even_mode ? 1 : 4;
    for (int t = 0; t < n[3]; t++)
        for (int z = 0; z < n[2]; z++)
            for (int y = 0; y < n[1]; y++)
                for (int xt = 0; xt < xlimit; xt++) {
                    const int x = scalar_even_mode ? 2 * xt + ((y + z + t) & 1) : xt;
                    const int parity = scalar_even_mode ? 0 : (x + y + z + t) % 2;
                    const int xs[4] = {x, y, z, t};
                    int d[4];
                    for (int k = 0; k < 4; k++) d[k] = xs[d0123[k]];
                    const int idxh = (int)snum_h(*L, d);
                    for (int dir = 0; dir < dirlimit; dir++)
                        single_element_rw(idxh, parity, L->map[dir], datastruct, swap, fp);
                }
}

int ostp_set_lattice_host(int nd0, int nd1, int nd2, int nd3, int par_temp) {
    if (nd0 <= 0 || nd1 <= 0 || nd2 <= 0 || nd3 <= 0 || (nd0 & 1))
        return fail("ostp_set_lattice_host: invalid lattice %dx%dx%dx%d", nd0, nd1, nd2, nd3) ? 0 : 1;
    Lattice L{};
    L.nd[0] = nd0;
    L.nd[1] = nd1;
    L.nd[2] = nd2;
    L.nd[3] = nd3;
    L.sizeh = (long)nd0 * nd1 * nd2 * nd3 / 2;
    L.par_temp = par_temp ? 1 : 0;
    for (int i = 0; i < 4; i++) L.map[i] = i;
    ctx().host_lat = L;
    ctx().host_lat_set = true;
    return 0;
}

// ------------------------------------------------------------ ASCII conf
// io.c:35-108: header "nx ny nz nt beta mass nflavours conf_id", then every
// matrix (row 2 rebuilt) as three lines of "(re, im)  "
int print_su3_soa_ASCII(const global_su3_soa* conf_rw, const char* nomefile, int conf_id_iter, const double_soa* back_phases) {
    const char* nm = "print_su3_soa_ASCII";
    const Lattice* L = io_lattice(nm);
    if (!L) return 2;
    const size_t sh = L->sizeh, stride = 144 * sh;
    HostView v(conf_rw, 8 * stride, true, nm);
    if (!v.ok) return 2;
    std::vector<double> ph;
    if (back_phases) {
        HostView pv(back_phases, 8 * 8 * sh, true, nm);
        if (!pv.ok) return 2;
        ph.assign((const double*)pv.p, (const double*)pv.p + 8 * sh);
        printf("This function works correctly if the field passed as argument contains\n");
        printf("ONLY STAGGERED PHASES\n");
    }
    FILE* fp = fopen(nomefile, "w");
    if (!fp) {
        printf("ERROR, %s unreadable.\n", nomefile);
        return 2;
    }
    int n[4];
    phys_extents(*L, n);
    const double fm = (alloc_info.NDiffFlavs != 0 && fermions_parameters) ? fermions_parameters[0].ferm_mass : 0.0;
    fprintf(fp, "%d %d %d %d %lf %lf %d %d\n", n[0], n[1], n[2], n[3], act_params.beta, fm, alloc_info.NDiffFlavs,
            conf_id_iter);
    const long expected = (long)n[0] * n[1] * n[2] * n[3] * 3 / 2;
    long minuses = 0;
    for (int q = 0; q < 8; q++)
        for (size_t i = 0; i < sh; i++) {
            cd m[3][3];
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) m[r][c] = at(v.p, stride, sh, q, r, c, i);
            rebuild3(m);
            int sg = 1;
            if (back_phases && fabs(ph[(size_t)q * sh + i]) > 0.001) {
                sg = -1;
                minuses++;
            }
            for (int r = 0; r < 3; r++) {
                for (int c = 0; c < 3; c++) fprintf(fp, "(%.18lf, %.18lf)  ", sg * m[r][c].real(), sg * m[r][c].imag());
                fprintf(fp, "\n");
            }
            fprintf(fp, "\n");
        }
    fclose(fp);
    if (back_phases) {
        printf("Written %ld matrices with -1 determinant, of expected %ld.\n", minuses, expected);
        if (minuses != expected) {
            printf("WRONG NUMBER OF DET=-1 MATRICES.\n");
            return 1;
        }
    }
    return 0;
}

// io.c:111-207: all three rows from the file; a matrix with determinant
// near -1 (a field saved with its staggered phases) is negated
int read_su3_soa_ASCII(global_su3_soa* conf, const char* nomefile, int* conf_id_iter) {
    const char* nm = "read_su3_soa_ASCII";
    const Lattice* L = io_lattice(nm);
    if (!L) return 2;
    FILE* fp = fopen(nomefile, "r");
    if (!fp) {
        printf("Gauge configuration file %s not readable.\n", nomefile);
        *conf_id_iter = -1;
        return 1;
    }
    const size_t sh = L->sizeh, stride = 144 * sh;
    HostView v(conf, 8 * stride, false, nm);
    if (!v.ok) {
        fclose(fp);
        return 2;
    }
    int n[4], nt[4], nfl;
    double beta, mass;
    phys_extents(*L, n);
    auto bad = [&](int line) {
        printf("%s:%d, Error, not read expected number of entries.\n", __FILE__, line);
        fclose(fp);
        return 2;
    };
    if (fscanf(fp, "%d %d %d %d %lf %lf %d %d\n", &nt[0], &nt[1], &nt[2], &nt[3], &beta, &mass, &nfl, conf_id_iter) != 8)
        return bad(__LINE__);
    if (n[0] != nt[0] || n[1] != nt[1] || n[2] != nt[2] || n[3] != nt[3]) {
        printf("Error, configuration dimensions not compatible with code.\n");
        printf("Code dimensions: d0 %d,d1 %d,d2 %d, d3 %d\n", n[0], n[1], n[2], n[3]);
        printf("Conf dimensions: d0 %d,d1 %d,d2 %d, d3 %d\n", nt[0], nt[1], nt[2], nt[3]);
        fclose(fp);
        return 2;
    }
    long minuses = 0;
    for (int q = 0; q < 8; q++)
        for (size_t i = 0; i < sh; i++) {
            cd m[3][3];
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                    double re, im;
                    if (fscanf(fp, c < 2 ? "(%lf, %lf) " : "(%lf, %lf)\n", &re, &im) != 2) return bad(__LINE__);
                    m[r][c] = cd(re, im);
                }
            const double det = det_re(m);
            const double sg = fabs(1 + det) < 0.005 ? -1.0 : 1.0;
            if (sg < 0) minuses++;
            else if (fabs(fabs(det) - 1.0) > 0.005)
                printf("Non unitary matrix read, Det M =  %.18lf\n", det);
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) at(v.p, stride, sh, q, r, c, i) = sg * m[r][c];
        }
    fclose(fp);
    if (minuses != 0) {
        printf("Conf read has some matrices with determinant = -1.\n");
        printf("The count of matrices with determinant = -1 is %ld (expected %ld for staggered phases).\n", minuses,
               (long)n[0] * n[1] * n[2] * n[3] * 3 / 2);
    }
    return v.flush(8 * stride, nm) ? 0 : 2;
}

// ------------------------------------------------------------ ILDG binary
// io.c:415-547: records ildg-format (XML), MD_traj, input-file,
// ildg-binary-data (64-bit big-endian, row 2 rebuilt), ildg-data-lfn (empty)
int print_su3_soa_ildg_binary(const global_su3_soa* conf, const char* nomefile, int conf_id_iter) {
    const char* nm = "print_su3_soa_ildg_binary";
    const Lattice* L = io_lattice(nm);
    if (!L) return 2;
    const size_t sh = L->sizeh, stride = 144 * sh;
    HostView v(conf, 8 * stride, true, nm);
    if (!v.ok) return 2;
    printf("Writing ildg conf...\n");
    FILE* fp = fopen(nomefile, "w");
    if (!fp) {
        printf("ERROR, %s unreadable.\n", nomefile);
        return 2;
    }
    int n[4];
    phys_extents(*L, n);
    char fmt[1000];
    snprintf(fmt, sizeof fmt,
             "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n"
             "            <ildgFormat xmlns=\"http://www.lqcd.org/ildg\"\n"
             "            xmlns:xsi=\"http://www.w3.org/2001/XMLSchema-instance\"\n"
             "            xsi:schemaLocation=\"http://www.lqcd.org/ildg filefmt.xsd\">\n"
             "            <version>1.0</version>\n"
             "            <field>su3gauge</field>\n"
             "            <precision>64</precision>\n"
             "            <lx>%d</lx>\n"
             "            <ly>%d</ly>\n"
             "            <lz>%d</lz>\n"
             "            <lt>%d</lt>\n"
             "            </ildgFormat>",
             n[0], n[1], n[2], n[3]);
    char traj[32];
    snprintf(traj, sizeof traj, "%d", conf_id_iter);
    bool ok = write_record(fp, "ildg-format", fmt, strlen(fmt)) && write_record(fp, "MD_traj", traj, strlen(traj)) &&
              write_record(fp, "input-file", input_file_str, strnlen(input_file_str, sizeof input_file_str));
    // the binary data record
    const uint64_t len = (uint64_t)n[0] * n[1] * n[2] * n[3] * 4 * 3 * 3 * 2 * 8;
    LimeHeader h;
    memset(&h, 0, sizeof h);
    h.magic_no = big_endian<uint32_t>(0x456789ab);
    h.version = big_endian<uint16_t>(1);
    h.data_length = big_endian<uint64_t>(len);
    strncpy(h.type, "ildg-binary-data", sizeof h.type - 1);
    ok = ok && fwrite(&h, sizeof h, 1, fp) == 1;
    std::vector<double> row(18);
    ildg_sites(*L, [&](int l, size_t i) {
        if (!ok) return;
        cd m[3][3];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) m[r][c] = at(v.p, stride, sh, l, r, c, i);
        rebuild3(m);
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                row[2 * (3 * r + c)] = big_endian(m[r][c].real());
                row[2 * (3 * r + c) + 1] = big_endian(m[r][c].imag());
            }
        ok = fwrite(row.data(), sizeof(double), 18, fp) == 18;
    });
    ok = ok && write_record(fp, "ildg-data-lfn", nullptr, 0);
    fclose(fp);
    if (!ok) fail("%s: writing %s failed", nm, nomefile);
    return ok ? 0 : 2;
}

// io.c:257-412: records until ildg-format and ildg-binary-data are found;
// extents from the XML, conf_id from MD_traj (1 when absent); row 2 of every
// matrix rebuilt from rows 0 and 1 as the reference does
int read_su3_soa_ildg_binary(global_su3_soa* conf, const char* nomefile, int* conf_id_iter) {
    const char* nm = "read_su3_soa_ildg_binary";
    const Lattice* L = io_lattice(nm);
    if (!L) return 2;
    *conf_id_iter = 1000;
    FILE* fg = fopen(nomefile, "r");
    if (!fg) {
        printf("Gauge configuration ILDG file %s not readable.\n", nomefile);
        *conf_id_iter = -1;
        return 1;
    }
    auto bad = [&](const char* why) {
        printf("Error in reading file %s: %s\n", nomefile, why);
        fclose(fg);
        return 2;
    };
    long fmt_pos = -1, bin_pos = -1, traj_pos = -1;
    uint64_t fmt_len = 0, bin_len = 0;
    for (int k = 0; k < 10 && (fmt_pos < 0 || bin_pos < 0); k++) {
        LimeHeader h;
        if (fread(&h, sizeof h, 1, fg) != 1) return bad("record header");
        h.type[sizeof h.type - 1] = 0;
        const uint64_t dl = big_endian(h.data_length);
        const long pos = ftell(fg);
        if (!strcmp(h.type, "ildg-format")) fmt_pos = pos, fmt_len = dl;
        if (!strcmp(h.type, "ildg-binary-data")) bin_pos = pos, bin_len = dl;
        if (!strcmp(h.type, "MD_traj")) traj_pos = pos;
        const uint64_t skip = dl + (dl % 8 == 0 ? 0 : 8 - dl % 8);
        if (fseek(fg, (long)skip, SEEK_CUR) != 0) return bad("record data");
    }
    if (fmt_pos < 0 || bin_pos < 0) return bad("ildg-format or ildg-binary-data record missing");
    if (fmt_len >= 4096) return bad("ildg-format record too long");
    char xml[4096];
    fseek(fg, fmt_pos, SEEK_SET);
    if (fread(xml, 1, fmt_len, fg) != fmt_len) return bad("ildg-format");
    xml[fmt_len] = 0;
    int nr[4];
    const char* tags[4] = {"<lx>", "<ly>", "<lz>", "<lt>"};
    for (int k = 0; k < 4; k++) {
        const char* s = strstr(xml, tags[k]);
        if (!s || sscanf(s + 4, " %d", &nr[k]) != 1) return bad("lx, ly, lz or lt not found in \"ildg-format\"");
    }
    int n[4];
    phys_extents(*L, n);
    if (n[0] != nr[0] || n[1] != nr[1] || n[2] != nr[2] || n[3] != nr[3]) {
        printf("Error, Configuration dimensions not compatible with input file.\n");
        fclose(fg);
        return 2;
    }
    if (traj_pos >= 0) {
        fseek(fg, traj_pos, SEEK_SET);
        if (fscanf(fg, "%d", conf_id_iter) != 1) return bad("MD_traj");
    } else {
        *conf_id_iter = 1;
    }
    const uint64_t want = (uint64_t)n[0] * n[1] * n[2] * n[3] * 4 * 18 * 8;
    if (bin_len < want) return bad("ildg-binary-data shorter than the lattice");
    const size_t sh = L->sizeh, stride = 144 * sh;
    HostView v(conf, 8 * stride, false, nm);
    if (!v.ok) {
        fclose(fg);
        return 2;
    }
    fseek(fg, bin_pos, SEEK_SET);
    std::vector<double> row(18);
    bool ok = true;
    ildg_sites(*L, [&](int l, size_t i) {
        if (!ok) return;
        ok = fread(row.data(), sizeof(double), 18, fg) == 18;
        if (!ok) return;
        cd m[3][3];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++)
                m[r][c] = cd(big_endian(row[2 * (3 * r + c)]), big_endian(row[2 * (3 * r + c) + 1]));
        rebuild3(m);
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) at(v.p, stride, sh, l, r, c, i) = m[r][c];
    });
    fclose(fg);
    if (!ok) return fail("%s: %s ends inside ildg-binary-data", nm, nomefile) ? 2 : 2;
    return v.flush(8 * stride, nm) ? 0 : 2;
}

// ------------------------------------------------------------ ASCII vec3
// dbgtools.c:1193-1236: one "re\tim" line per colour component, sites in order
void print_vec3_soa(const vec3_soa* fermion, const char* nomefile) {
    const char* nm = "print_vec3_soa";
    const Lattice* L = io_lattice(nm);
    if (!L) return;
    const size_t sh = L->sizeh;
    HostView v(fermion, 48 * sh, true, nm);
    if (!v.ok) return;
    FILE* fp = fopen(nomefile, "w");
    if (!fp) {
        fail("%s: cannot open %s", nm, nomefile);
        return;
    }
    const cd* c = (const cd*)v.p;
    for (size_t i = 0; i < sh; i++)
        for (int k = 0; k < 3; k++) fprintf(fp, "%.18lf\t%.18lf\n", c[k * sh + i].real(), c[k * sh + i].imag());
    fclose(fp);
}

int read_vec3_soa(vec3_soa* fermion, const char* nomefile) {
    const char* nm = "read_vec3_soa";
    const Lattice* L = io_lattice(nm);
    if (!L) return 2;
    FILE* fp = fopen(nomefile, "r");
    if (!fp) {
        printf("vec3_soa File %s not found.\n", nomefile);
        return 1;
    }
    const size_t sh = L->sizeh;
    HostView v(fermion, 48 * sh, false, nm);
    if (!v.ok) {
        fclose(fp);
        return 2;
    }
    cd* c = (cd*)v.p;
    for (size_t i = 0; i < sh; i++)
        for (int k = 0; k < 3; k++) {
            double re, im;
            if (fscanf(fp, "%lf\t%lf\n", &re, &im) != 2) {
                fclose(fp);
                return fail("%s: %s: not read expected number of entries at site %zu", nm, nomefile, i) ? 2 : 2;
            }
            c[k * sh + i] = cd(re, im);
        }
    fclose(fp);
    return v.flush(48 * sh, nm) ? 0 : 2;
}

// ------------------------------------------------ wrappers on su3_soa[8]
// save_conf_wrapper / read_conf_wrapper (io.h:51-150) for the library's
// su3_soa arrays (device or registered host addresses, K plane skipped)
int ostp_save_conf(const su3_soa* conf, const char* nomefile, int conf_id_iter, int use_ildg) {
    const char* nm = "ostp_save_conf";
    if (!require_init(nm)) return 2;
    if (ctx().host_lat_set) return fail("%s: ostp_set_lattice_host is in effect; ostp_init first", nm) ? 2 : 2;
    const Lattice& L = ctx().lat;
    const size_t sh = L.sizeh, gstride = 144 * sh;
    const char* d = (const char*)resolve(conf, 8 * L.su3_bytes(), nm);
    if (!d) return 2;
    std::vector<char> host(8 * gstride);
    for (int l = 0; l < 8; l++)
        if (ostp_memcpy_dtoh(host.data() + l * gstride, d + l * L.su3_bytes(), gstride) != 0) return 2;
    return use_ildg ? print_su3_soa_ildg_binary((const global_su3_soa*)host.data(), nomefile, conf_id_iter)
                    : print_su3_soa_ASCII((const global_su3_soa*)host.data(), nomefile, conf_id_iter, nullptr);
}

int ostp_read_conf(su3_soa* conf, const char* nomefile, int* conf_id_iter, int use_ildg) {
    const char* nm = "ostp_read_conf";
    if (!require_init(nm)) return 2;
    if (ctx().host_lat_set) return fail("%s: ostp_set_lattice_host is in effect; ostp_init first", nm) ? 2 : 2;
    const Lattice& L = ctx().lat;
    const size_t sh = L.sizeh, gstride = 144 * sh;
    char* d = (char*)resolve(conf, 8 * L.su3_bytes(), nm);
    if (!d) return 2;
    std::vector<char> host(8 * gstride);
    const int e = use_ildg ? read_su3_soa_ildg_binary((global_su3_soa*)host.data(), nomefile, conf_id_iter)
                           : read_su3_soa_ASCII((global_su3_soa*)host.data(), nomefile, conf_id_iter);
    if (e) return e;
    for (int l = 0; l < 8; l++)
        if (ostp_memcpy_htod(d + l * L.su3_bytes(), host.data() + l * gstride, gstride) != 0) return 2;
    // a registered host array gets the same contents, as after the reference's
    // host-side read
    if (ostp_acc_is_present(conf, 8 * L.su3_bytes()) && (void*)d != (void*)conf)
        for (int l = 0; l < 8; l++) memcpy((char*)conf + l * L.su3_bytes(), host.data() + l * gstride, gstride);
    return 0;
}

}  // extern "C"
