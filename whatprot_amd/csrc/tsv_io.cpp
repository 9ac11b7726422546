// The reference's text formats, read and written in parallel: radiometries
// (io/radiometries-io.cc:33-81), dye tracks (io/dye-tracks-io.cc:32-103) and
// results (io/scored-classifications-io.cc:54-68).
//
// A file is read whole (several threads, pread), its records (one per line, as the reference writes
// them) split into line-aligned chunks, and each chunk parsed by its own
// thread with std::from_chars, which rounds decimal text to the nearest
// double exactly as the reference's operator>> (strtod) does. A file whose
// lines do not hold one record each is parsed as one whitespace-separated
// token stream instead, as operator>> would. Writers format with
// std::to_chars, whose precision-17 general form is printf's "%.17g", the
// output of operator<< after setprecision(17), so files are byte-identical
// to the reference's; chunks are formatted in parallel and each written at
// its offset by its own thread (pwrite).
#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include "wp_internal.h"

namespace wp {
namespace {

struct Buffer {
    std::vector<char> data;
};

int threads_for(size_t bytes);
template <typename F>
void parallel(int k, F f);

// The whole file, read by several threads with pread.
int read_file(const char* path, Buffer* b) {
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0) {
        set_error(std::string("cannot open ") + path);
        return WP_ERR_ARG;
    }
    struct stat st;
    if (::fstat(fd, &st) != 0 || st.st_size < 0) {
        ::close(fd);
        set_error(std::string("cannot size ") + path);
        return WP_ERR_ARG;
    }
    const size_t n = (size_t)st.st_size;
    b->data.resize(n);
    const int k = threads_for(n);
    std::vector<int> ok(k, 1);
    parallel(k, [&](int i) {
        size_t a = n * i / k, e = n * (i + 1) / k;
        int good = 1;
        while (a < e) {
            const ssize_t got = ::pread(fd, b->data.data() + a, e - a, (off_t)a);
            if (got <= 0) {
                good = 0;
                break;
            }
            a += (size_t)got;
        }
        ok[i] = good;
    });
    ::close(fd);
    for (int v : ok)
        if (!v) {
            set_error(std::string("short read of ") + path);
            return WP_ERR_ARG;
        }
    return WP_OK;
}

// head, then the pieces in order, each piece written by its own thread at
// its offset.
int write_file(const char* path, const std::string& head, const std::vector<std::string>& pieces) {
    const int fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) {
        set_error(std::string("cannot create ") + path);
        return WP_ERR_ARG;
    }
    const int k = (int)pieces.size() + 1;
    std::vector<size_t> off(k + 1, 0);
    off[1] = head.size();
    for (int i = 1; i < k; i++) off[i + 1] = off[i] + pieces[i - 1].size();
    std::vector<int> ok(k, 1);
    parallel(k, [&](int i) {
        const std::string& s = i == 0 ? head : pieces[i - 1];
        size_t a = 0;
        int good = 1;
        while (a < s.size()) {
            const ssize_t put = ::pwrite(fd, s.data() + a, s.size() - a, (off_t)(off[i] + a));
            if (put <= 0) {
                good = 0;
                break;
            }
            a += (size_t)put;
        }
        ok[i] = good;
    });
    bool good = ::close(fd) == 0;
    for (int v : ok) good = good && v;
    if (!good) {
        set_error(std::string("cannot write ") + path);
        return WP_ERR_ARG;
    }
    return WP_OK;
}

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

// Whitespace-separated tokens of [p, end).
struct Cursor {
    const char* p;
    const char* end;
    void skip() {
        while (p < end && is_space(*p)) p++;
    }
    template <typename V>
    bool next(V* v) {
        skip();
        if (p >= end) return false;
        const char* q = p;
        if (*q == '+' && q + 1 < end && !is_space(q[1])) q++;  // operator>> accepts a leading '+', from_chars does not
        auto r = std::from_chars(q, end, *v);
        if (r.ec != std::errc() || (r.ptr < end && !is_space(*r.ptr))) return false;
        p = r.ptr;
        return true;
    }
};

int threads_for(size_t bytes) {
    unsigned hw = std::thread::hardware_concurrency();
    int n = (int)std::min<unsigned>(hw ? hw : 4, 32);
    return std::max(1, std::min<int>(n, (int)(bytes >> 20) + 1));  // about 1 MB or more per thread
}

// [begin, end) split into k line-aligned pieces.
std::vector<const char*> line_chunks(const char* begin, const char* end, int k) {
    std::vector<const char*> cut{begin};
    const size_t len = (size_t)(end - begin);
    for (int i = 1; i < k; i++) {
        const char* p = begin + len * i / k;
        if (p < cut.back()) p = cut.back();
        while (p < end && *p != '\n') p++;
        if (p < end) p++;
        cut.push_back(p);
    }
    cut.push_back(end);
    return cut;
}

template <typename F>
void parallel(int k, F f) {
    std::vector<std::thread> th;
    for (int i = 1; i < k; i++) th.emplace_back(f, i);
    f(0);
    for (auto& t : th) t.join();
}

// Non-blank lines of [a, b) as (start, end) pairs.
void lines_of(const char* a, const char* b, std::vector<std::pair<const char*, const char*>>* out) {
    std::vector<std::pair<const char*, const char*>> v;  // local: see the writers
    while (a < b) {
        const char* e = (const char*)std::memchr(a, '\n', (size_t)(b - a));
        if (!e) e = b;
        const char* q = a;
        while (q < e && is_space(*q)) q++;
        if (q < e) v.push_back({a, e});
        a = e + 1;
    }
    *out = std::move(v);
}

}  // namespace
}  // namespace wp

extern "C" {

int wp_read_radiometries(const char* path, const wp_model* seq_model, wp_radiometries* out) {
    if (!path || !out) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    std::memset(out, 0, sizeof(*out));
    wp::Buffer buf;
    int rc = wp::read_file(path, &buf);
    if (rc) return rc;
    const char* begin = buf.data.data();
    const char* end = begin + buf.data.size();
    wp::Cursor cur{begin, end};
    unsigned int T = 0, C = 0, N = 0;
    if (!cur.next(&T) || !cur.next(&C) || !cur.next(&N)) {
        wp::set_error("radiometries: missing num_timesteps, num_channels or num_radiometries");
        return WP_ERR_ARG;
    }
    if (seq_model && seq_model->num_channels != (int)C) {
        wp::set_error("radiometries: the file's channels differ from the model's");
        return WP_ERR_ARG;
    }
    if (T < 1 || C < 1 || C > WP_MAX_CHANNELS) {
        wp::set_error("radiometries: bad shape");
        return WP_ERR_ARG;
    }
    const size_t per = (size_t)T * C;
    double* v = (double*)std::malloc(sizeof(double) * std::max<size_t>(1, per * N));
    if (!v) {
        wp::set_error("radiometries: out of memory");
        return WP_ERR_NOMEM;
    }
    double mu[WP_MAX_CHANNELS];
    for (unsigned c = 0; c < C; c++) mu[c] = seq_model ? seq_model->channels[c].mu : 1.0;
    // one record per line: parse the lines in parallel
    const int k = wp::threads_for((size_t)(end - cur.p));
    std::vector<const char*> cut = wp::line_chunks(cur.p, end, k);
    std::vector<std::vector<std::pair<const char*, const char*>>> lines(k);
    wp::parallel(k, [&](int i) { wp::lines_of(cut[i], cut[i + 1], &lines[i]); });
    size_t total = 0;
    std::vector<size_t> base(k + 1, 0);
    for (int i = 0; i < k; i++) base[i + 1] = (total += lines[i].size());
    std::vector<int> ok(k, 1);
    if (total == N) {
        wp::parallel(k, [&](int i) {
            size_t r = base[i];
            int good = 1;
            for (const auto& ln : lines[i]) {
                wp::Cursor c{ln.first, ln.second};
                double* row = v + r * per;
                for (size_t j = 0; j < per && good; j++) {
                    double x;
                    if (!c.next(&x)) good = 0;
                    else row[j] = x / mu[j % C];
                }
                c.skip();
                if (c.p != c.end) good = 0;
                if (!good) break;
                r++;
            }
            ok[i] = good;
        });
    }
    bool fast = total == N;
    for (int i = 0; i < k; i++) fast = fast && ok[i];
    if (!fast) {  // not one record per line: one token stream, as operator>> reads it
        wp::Cursor c{cur.p, end};
        for (size_t j = 0; j < per * N; j++) {
            double x;
            if (!c.next(&x)) {
                std::free(v);
                wp::set_error("radiometries: fewer values than num_radiometries x num_timesteps x num_channels");
                return WP_ERR_ARG;
            }
            v[j] = x / mu[j % C];
        }
    }
    out->num_timesteps = (int)T;
    out->num_channels = (int)C;
    out->num_reads = (int)N;
    out->values = v;
    return WP_OK;
}

void wp_free_radiometries(wp_radiometries* r) {
    if (!r) return;
    std::free(r->values);
    r->values = nullptr;
}

int wp_write_radiometries(const char* path, const wp_model* seq_model, int T, int C, int N, const double* values) {
    if (!path || T < 1 || C < 1 || C > WP_MAX_CHANNELS || N < 0 || (N > 0 && !values) ||
        (seq_model && seq_model->num_channels != C)) {
        wp::set_error("bad radiometries arguments");
        return WP_ERR_ARG;
    }
    const std::string head = std::to_string(T) + "\n" + std::to_string(C) + "\n" + std::to_string(N) + "\n";
    double mu[WP_MAX_CHANNELS];
    for (int c = 0; c < C; c++) mu[c] = seq_model ? seq_model->channels[c].mu : 1.0;
    const size_t per = (size_t)T * C;
    const int k = std::max(1, std::min(wp::threads_for(per * (size_t)N * 24), N));
    std::vector<std::string> text(k);
    // each thread builds its own string (neighbouring headers in one vector
    // would share cache lines on every append) and hands it over at the end
    wp::parallel(k, [&](int i) {
        const size_t r0 = (size_t)N * i / k, r1 = (size_t)N * (i + 1) / k;
        std::string s;
        s.reserve((r1 - r0) * per * 24);
        char tmp[64];
        for (size_t r = r0; r < r1; r++) {
            for (size_t j = 0; j < per; j++) {
                if (j) s.push_back('\t');
                auto res = std::to_chars(tmp, tmp + sizeof(tmp), mu[j % C] * values[r * per + j],
                                         std::chars_format::general, 17);
                s.append(tmp, res.ptr);
            }
            s.push_back('\n');
        }
        text[i] = std::move(s);
    });
    return wp::write_file(path, head, text);
}

int wp_read_dye_tracks(const char* path, wp_dye_tracks* out) {
    if (!path || !out) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    std::memset(out, 0, sizeof(*out));
    wp::Buffer buf;
    int rc = wp::read_file(path, &buf);
    if (rc) return rc;
    const char* begin = buf.data.data();
    const char* end = begin + buf.data.size();
    wp::Cursor cur{begin, end};
    unsigned int T = 0, C = 0, n = 0;
    if (!cur.next(&T) || !cur.next(&C) || !cur.next(&n) || T < 1 || C < 1 || C > WP_MAX_CHANNELS) {
        wp::set_error("dye tracks: missing or bad num_timesteps, num_channels or num_dye_tracks");
        return WP_ERR_ARG;
    }
    const size_t per = (size_t)T * C;
    struct Part {
        std::vector<short> counts;
        std::vector<int> nsrc, ids, cnts, hits;
        bool ok = true;
    };
    // one track per line, parsed in parallel into per-chunk parts
    const int k = wp::threads_for((size_t)(end - cur.p));
    std::vector<const char*> cut = wp::line_chunks(cur.p, end, k);
    std::vector<Part> parts(k);
    auto parse_track = [&](wp::Cursor& c, Part& pt) {
        for (size_t j = 0; j < per; j++) {
            short x;
            if (!c.next(&x)) return false;
            pt.counts.push_back(x);
        }
        int ns;
        if (!c.next(&ns) || ns < 0) return false;
        pt.nsrc.push_back(ns);
        for (int j = 0; j < ns; j++) {
            int a, b, h;
            if (!c.next(&a) || !c.next(&b) || !c.next(&h)) return false;
            pt.ids.push_back(a);
            pt.cnts.push_back(b);
            pt.hits.push_back(h);
        }
        return true;
    };
    wp::parallel(k, [&](int i) {  // thread-local part, handed over at the end
        std::vector<std::pair<const char*, const char*>> lines;
        wp::lines_of(cut[i], cut[i + 1], &lines);
        Part pt;
        pt.counts.reserve(lines.size() * per);
        pt.nsrc.reserve(lines.size());
        pt.ids.reserve(lines.size() * 2);
        pt.cnts.reserve(lines.size() * 2);
        pt.hits.reserve(lines.size() * 2);
        for (const auto& ln : lines) {
            wp::Cursor c{ln.first, ln.second};
            if (!parse_track(c, pt)) {
                pt.ok = false;
                break;
            }
            c.skip();
            if (c.p != c.end) {
                pt.ok = false;
                break;
            }
        }
        parts[i] = std::move(pt);
    });
    size_t tracks = 0;
    bool fast = true;
    for (const Part& pt : parts) {
        tracks += pt.nsrc.size();
        fast = fast && pt.ok;
    }
    if (!fast || tracks != n) {  // not one track per line: one token stream
        parts.assign(1, Part());
        wp::Cursor c{cur.p, end};
        for (unsigned i = 0; i < n; i++) {
            if (!parse_track(c, parts[0])) {
                wp::set_error("dye tracks: fewer or malformed entries than num_dye_tracks");
                return WP_ERR_ARG;
            }
        }
    }
    size_t ns = 0;
    for (const Part& pt : parts) ns += pt.ids.size();
    out->counts = (short*)std::malloc(sizeof(short) * std::max<size_t>(1, per * n));
    out->source_offsets = (int*)std::malloc(sizeof(int) * ((size_t)n + 1));
    out->source_ids = (int*)std::malloc(sizeof(int) * std::max<size_t>(1, ns));
    out->source_counts = (int*)std::malloc(sizeof(int) * std::max<size_t>(1, ns));
    out->source_hits = (int*)std::malloc(sizeof(int) * std::max<size_t>(1, ns));
    if (!out->counts || !out->source_offsets || !out->source_ids || !out->source_counts || !out->source_hits) {
        wp_free_dye_tracks(out);
        wp::set_error("dye tracks: out of memory");
        return WP_ERR_NOMEM;
    }
    // each part copied into place by its own thread
    const int np = (int)parts.size();
    std::vector<size_t> tb(np + 1, 0), sb(np + 1, 0);
    for (int i = 0; i < np; i++) {
        tb[i + 1] = tb[i] + parts[i].nsrc.size();
        sb[i + 1] = sb[i] + parts[i].ids.size();
    }
    out->source_offsets[0] = 0;
    wp::parallel(np, [&](int i) {
        const Part& pt = parts[i];
        std::memcpy(out->counts + tb[i] * per, pt.counts.data(), sizeof(short) * pt.counts.size());
        int o = (int)sb[i];
        for (size_t j = 0; j < pt.nsrc.size(); j++) {
            o += pt.nsrc[j];
            out->source_offsets[tb[i] + j + 1] = o;
        }
        std::memcpy(out->source_ids + sb[i], pt.ids.data(), sizeof(int) * pt.ids.size());
        std::memcpy(out->source_counts + sb[i], pt.cnts.data(), sizeof(int) * pt.cnts.size());
        std::memcpy(out->source_hits + sb[i], pt.hits.data(), sizeof(int) * pt.hits.size());
        std::vector<short>().swap(parts[i].counts);  // free in parallel too
    });
    out->num_timesteps = (int)T;
    out->num_channels = (int)C;
    out->num_dye_tracks = (int)n;
    out->num_sources = (int)ns;
    return WP_OK;
}

void wp_free_dye_tracks(wp_dye_tracks* t) {
    if (!t) return;
    std::free(t->counts);
    std::free(t->source_offsets);
    std::free(t->source_ids);
    std::free(t->source_counts);
    std::free(t->source_hits);
    t->counts = nullptr;
    t->source_offsets = t->source_ids = t->source_counts = t->source_hits = nullptr;
}

int wp_write_dye_tracks(const char* path, int T, int C, int n, const short* counts, const int* source_offsets,
                        const int* source_ids, const int* source_counts, const int* source_hits) {
    if (!path || T < 1 || C < 1 || n < 0 ||
        (n > 0 && (!counts || !source_offsets || !source_ids || !source_counts || !source_hits))) {
        wp::set_error("bad dye-track arguments");
        return WP_ERR_ARG;
    }
    const std::string head = std::to_string(T) + "\n" + std::to_string(C) + "\n" + std::to_string(n) + "\n";
    const size_t per = (size_t)T * C;
    const int k = std::max(1, std::min(wp::threads_for((size_t)n * (per * 2 + 12)), std::max(n, 1)));
    std::vector<std::string> text(k);
    wp::parallel(k, [&](int i) {  // thread-local string, handed over at the end
        const size_t r0 = (size_t)n * i / k, r1 = (size_t)n * (i + 1) / k;
        std::string s;
        s.reserve((r1 - r0) * (per * 2 + 16));
        char tmp[16];
        auto put = [&](int x) {
            auto res = std::to_chars(tmp, tmp + sizeof(tmp), x);
            s.append(tmp, res.ptr);
        };
        for (size_t r = r0; r < r1; r++) {
            for (size_t j = 0; j < per; j++) {
                put(counts[r * per + j]);
                s.push_back('\t');
            }
            const int a = source_offsets[r], b = source_offsets[r + 1];
            put(b - a);
            s.push_back('\t');
            for (int j = a; j < b; j++) {
                put(source_ids[j]);
                s.push_back('\t');
                put(source_counts[j]);
                s.push_back('\t');
                put(source_hits[j]);
                if (j < b - 1) s.push_back('\t');
            }
            s.push_back('\n');
        }
        text[i] = std::move(s);
    });
    return wp::write_file(path, head, text);
}

int wp_write_scored_classifications(const char* path, int n, const int* ids, const double* scores) {
    if (!path || n < 0 || (n > 0 && (!ids || !scores))) {
        wp::set_error("bad results arguments");
        return WP_ERR_ARG;
    }
    std::string s = "radmat_iz,best_pep_iz,best_pep_score\n";
    s.reserve(s.size() + (size_t)n * 36);
    char tmp[64];
    for (int i = 0; i < n; i++) {
        auto r = std::to_chars(tmp, tmp + sizeof(tmp), i);
        s.append(tmp, r.ptr);
        s.push_back(',');
        r = std::to_chars(tmp, tmp + sizeof(tmp), ids[i]);
        s.append(tmp, r.ptr);
        s.push_back(',');
        r = std::to_chars(tmp, tmp + sizeof(tmp), scores[i], std::chars_format::general, 17);
        s.append(tmp, r.ptr);
        s.push_back('\n');
    }
    return wp::write_file(path, s, {});
}

}  // extern "C"
