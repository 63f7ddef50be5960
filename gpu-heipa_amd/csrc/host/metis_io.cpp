// METIS graph reader/writer and partition file I/O.
//
// Reader semantics follow the reference reader (src/datastructures/
// host_graph.h:154-247): header "n m [fmt]", '%' comment lines, 1-based
// neighbour ids, m in the header counts undirected edges and the CSR stores
// both directions, g_weight = sum of vertex weights, uniform weights are not
// stored.  Deliberate difference: the reference copies up to three raw fmt
// characters left-aligned, so "1" and "11" are read as "100"/"110"; here fmt is
// right-aligned as the METIS manual specifies ("1" == "001").  Both agree on
// every 3-character fmt, which is what the generators write.  Errors are
// returned as messages instead of exit(EXIT_FAILURE).
//
// The partition writer produces the reference's output format: one block id
// per line, '\n' terminated, n lines (src/utility/util.h:288-329).
#include "host_graph.hpp"

#include <cerrno>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <fcntl.h>
#include <fstream>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

namespace heipa {
namespace {

struct Cursor {
    const char *p;
    const char *end;
    bool eof() const { return p >= end; }
    void skip_spaces() { while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p; }
    void skip_line() {
        while (p < end && *p != '\n') ++p;
        if (p < end) ++p;
    }
    void skip_comments() {
        while (p < end && *p == '%') skip_line();
    }
    bool at_eol() const { return p >= end || *p == '\n'; }
    bool read_u64(u64 &x) {
        skip_spaces();
        if (p >= end || *p < '0' || *p > '9') return false;
        x = 0;
        while (p < end && *p >= '0' && *p <= '9') {
            x = x * 10 + (u64) (*p - '0');
            ++p;
        }
        return true;
    }
};

} // namespace

std::string read_metis(const std::string &path, HostGraph &g) {
    int fd = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
    if (fd < 0) return "File " + path + " does not exist!";
    struct stat st{};
    if (fstat(fd, &st) != 0) {
        ::close(fd);
        return "cannot stat " + path;
    }
    const size_t size = (size_t) st.st_size;
    const char *data = nullptr;
    void *addr = nullptr;
    if (size > 0) {
        addr = ::mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (addr == MAP_FAILED) {
            ::close(fd);
            return "cannot mmap " + path;
        }
        data = (const char *) addr;
    }
    std::string err;
    Cursor c{data, data + size};
    auto fail = [&](const std::string &msg) {
        if (addr) ::munmap(addr, size);
        ::close(fd);
        return msg;
    };
    c.skip_comments();
    u64 n = 0, m_und = 0;
    if (!c.read_u64(n) || !c.read_u64(m_und)) return fail("malformed METIS header in " + path);
    std::string fmt = "000";
    c.skip_spaces();
    if (!c.at_eol()) {
        u64 f = 0;
        const char *s = c.p;
        if (!c.read_u64(f)) return fail("malformed fmt field in " + path);
        std::string raw(s, c.p);
        if (raw.size() > 3) return fail("unsupported fmt '" + raw + "' in " + path);
        fmt = std::string(3 - raw.size(), '0') + raw;
        c.skip_spaces();
        if (!c.at_eol()) return fail("ncon field is not supported in " + path);
    }
    if (fmt[0] != '0') return fail("vertex sizes (fmt 1xx) are not supported in " + path);
    if (n >= (u64) UINT32_MAX || 2 * m_und >= (u64) UINT32_MAX) return fail("graph too large for u32 ids/offsets");
    const bool has_vw = fmt[1] == '1', has_ew = fmt[2] == '1';
    c.skip_line();

    g = HostGraph{};
    g.n = (u32) n;
    g.m = (u32) (2 * m_und);
    g.uniform_vertex_weights = !has_vw;
    g.uniform_edge_weights = !has_ew;
    g.xadj.assign(g.n + 1, 0);
    g.adjncy.resize(g.m);
    if (has_ew) g.adjwgt.resize(g.m);
    if (has_vw) g.vwgt.resize(g.n);
    u64 cur = 0;
    s64 gw = 0;
    for (u32 u = 0; u < g.n; ++u) {
        c.skip_comments();
        if (c.eof()) return fail("unexpected end of file at vertex " + std::to_string(u + 1));
        s64 vw = 1;
        if (has_vw) {
            u64 x;
            if (!c.read_u64(x)) return fail("missing vertex weight at vertex " + std::to_string(u + 1));
            vw = (s64) x;
            g.vwgt[u] = (s32) x;
        }
        gw += vw;
        while (true) {
            c.skip_spaces();
            if (c.at_eol()) break;
            u64 v;
            if (!c.read_u64(v)) return fail("malformed adjacency at vertex " + std::to_string(u + 1));
            if (v == 0 || v > n) return fail("neighbour id out of range at vertex " + std::to_string(u + 1));
            if (cur >= g.m) return fail("more edges than the header announces");
            g.adjncy[cur] = (u32) (v - 1);
            if (has_ew) {
                u64 w;
                if (!c.read_u64(w)) return fail("missing edge weight at vertex " + std::to_string(u + 1));
                g.adjwgt[cur] = (s32) w;
            }
            ++cur;
        }
        g.xadj[u + 1] = (u32) cur;
        if (!c.eof()) ++c.p; // newline
    }
    if (cur != g.m)
        return fail("Number of expected edges " + std::to_string(g.m) + " not equal to number edges " +
                    std::to_string(cur) + " found!");
    g.g_weight = gw;
    if (addr) ::munmap(addr, size);
    ::close(fd);
    return err;
}

std::string write_metis(const std::string &path, const HostGraph &g) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return "cannot open " + path;
    std::string buf;
    buf.reserve(1 << 20);
    char tmp[32];
    auto put = [&](u64 x) {
        auto r = std::to_chars(tmp, tmp + sizeof(tmp), x);
        buf.append(tmp, (size_t) (r.ptr - tmp));
    };
    auto flush = [&]() {
        if (buf.size() > (1u << 19)) {
            std::fwrite(buf.data(), 1, buf.size(), f);
            buf.clear();
        }
    };
    put(g.n);
    buf.push_back(' ');
    put(g.m / 2);
    const bool vw = !g.uniform_vertex_weights, ew = !g.uniform_edge_weights;
    if (vw || ew) {
        buf.append(" 0");
        buf.push_back(vw ? '1' : '0');
        buf.push_back(ew ? '1' : '0');
    }
    buf.push_back('\n');
    for (u32 u = 0; u < g.n; ++u) {
        bool first = true;
        if (vw) {
            put((u64) g.vwgt[u]);
            first = false;
        }
        for (u32 j = g.xadj[u]; j < g.xadj[u + 1]; ++j) {
            if (!first) buf.push_back(' ');
            first = false;
            put((u64) g.adjncy[j] + 1);
            if (ew) {
                buf.push_back(' ');
                put((u64) g.adjwgt[j]);
            }
        }
        buf.push_back('\n');
        flush();
    }
    std::fwrite(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    return {};
}

std::string write_partition(const std::string &path, const u32 *part, u32 n) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return "Error: Could not open " + path + " to write partition!";
    std::string buf;
    buf.reserve(1 << 20);
    char tmp[16];
    for (u32 u = 0; u < n; ++u) {
        auto r = std::to_chars(tmp, tmp + sizeof(tmp), part[u]);
        buf.append(tmp, (size_t) (r.ptr - tmp));
        buf.push_back('\n');
        if (buf.size() > (1u << 19)) {
            std::fwrite(buf.data(), 1, buf.size(), f);
            buf.clear();
        }
    }
    std::fwrite(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    return {};
}

std::string read_partition(const std::string &path, std::vector<u32> &part) {
    std::ifstream in(path);
    if (!in) return "Cannot open partition file " + path;
    part.clear();
    u64 x;
    while (in >> x) part.push_back((u32) x);
    return {};
}

} // namespace heipa
