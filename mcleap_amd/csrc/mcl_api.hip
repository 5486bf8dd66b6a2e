// mcl_api.hip -- host-mesh construction entry points and library info.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "mcl_internal.hpp"

struct mcl_hmesh {
    mcl::HostMesh m;
};

extern "C" int mcl_hmesh_build(const float* xy, int n, int periodic, mcl_hmesh_t** out) {
    if (!out || (n > 0 && !xy) || n < 0) return MCL_ERR_INPUT;
    if ((long long)n * 6 + 10 >= (1LL << 31)) return MCL_ERR_CAPACITY;
    mcl_hmesh* h = new mcl_hmesh();
    int rc = periodic ? mcl::build_torus(xy, n, h->m) : mcl::build_box(xy, n, h->m);
    if (rc != MCL_OK) { delete h; return rc; }
    *out = h;
    return MCL_OK;
}

// Ensemble of independent box triangulations as one mesh: every instance is
// built on its own (build_box, one instance per host thread, largest first),
// then copied into the union with all corners first (instance i: 4i..4i+3)
// and each instance's points, half-edges and triangles after the previous
// instance's.  Half-edge offsets stay even, so twins remain h ^ 1.
extern "C" int mcl_hmesh_build_batch(const float* xy, const int* counts, int ninst, int nthreads,
                                     mcl_hmesh_t** out) {
    if (!out || !counts || ninst <= 0) return MCL_ERR_INPUT;
    std::vector<long long> pstart(ninst + 1, 0);
    for (int i = 0; i < ninst; i++) {
        if (counts[i] < 0) return MCL_ERR_INPUT;
        pstart[i + 1] = pstart[i] + counts[i];
    }
    const long long npts = pstart[ninst];
    if (npts > 0 && !xy) return MCL_ERR_INPUT;
    if (npts + 4LL * ninst >= (1LL << 31) / 6) return MCL_ERR_CAPACITY;
    std::vector<mcl::HostMesh> part(ninst);
    std::vector<int> rc(ninst, MCL_OK);
    std::vector<int> order(ninst);
    for (int i = 0; i < ninst; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return counts[a] > counts[b]; });
    int nt = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
    nt = std::max(1, std::min(nt, ninst));
    std::atomic<int> next{0};
    auto build = [&]() {
        for (int k; (k = next.fetch_add(1)) < ninst;) {
            const int i = order[k];
            rc[i] = mcl::build_box(xy + 2 * pstart[i], counts[i], part[i]);
        }
    };
    std::vector<std::thread> pool;
    for (int w = 1; w < nt; w++) pool.emplace_back(build);
    build();
    for (auto& th : pool) th.join();
    for (int i = 0; i < ninst; i++)
        if (rc[i] != MCL_OK) return rc[i];
    mcl_hmesh* h = new mcl_hmesh();
    mcl::HostMesh& m = h->m;
    m.periodic = false;
    m.vstart.resize(ninst + 1);
    m.hestart.resize(ninst + 1);
    m.tstart.resize(ninst + 1);
    long long nv = 4LL * ninst, nhe = 0, ntri = 0;
    for (int i = 0; i < ninst; i++) {
        m.vstart[i] = (int)nv;
        m.hestart[i] = (int)nhe;
        m.tstart[i] = (int)ntri;
        nv += part[i].nv - 4;
        nhe += (long long)part[i].he.size() / 4;
        ntri += (long long)part[i].t.size();
        m.flips += part[i].flips;
    }
    if (nhe >= (1LL << 31)) { delete h; return MCL_ERR_CAPACITY; }
    m.vstart[ninst] = (int)nv;
    m.hestart[ninst] = (int)nhe;
    m.tstart[ninst] = (int)ntri;
    m.nv = (int)nv;
    m.pos.resize(2 * (size_t)nv);
    m.vrec.resize(2 * (size_t)nv);
    m.he.resize(4 * (size_t)nhe);
    m.t.resize((size_t)ntri);
    m.perm.resize((size_t)npts);
    next = 0;
    auto copy = [&]() {
        for (int i; (i = next.fetch_add(1)) < ninst;) {
            mcl::HostMesh& q = part[i];
            const int vs = m.vstart[i], hs = m.hestart[i], ts = m.tstart[i];
            auto vmap = [&](int v) { return v < 0 ? v : (v < 4 ? 4 * i + v : vs + v - 4); };
            for (int v = 0; v < q.nv; v++) {
                const int g = vmap(v);
                m.pos[2 * (size_t)g] = q.pos[2 * (size_t)v];
                m.pos[2 * (size_t)g + 1] = q.pos[2 * (size_t)v + 1];
                m.vrec[2 * (size_t)g] = g;
                m.vrec[2 * (size_t)g + 1] = q.vrec[2 * (size_t)v + 1] + hs;
            }
            const size_t nh = q.he.size() / 4;
            for (size_t e = 0; e < nh; e++) {
                int* o = &m.he[4 * ((size_t)hs + e)];
                o[0] = q.he[4 * e] + hs;
                o[1] = vmap(q.he[4 * e + 1]);
                o[2] = q.he[4 * e + 2] < 0 ? q.he[4 * e + 2] : q.he[4 * e + 2] + ts;
                o[3] = vmap(q.he[4 * e + 3]);
            }
            for (size_t k = 0; k < q.t.size(); k++) m.t[(size_t)ts + k] = q.t[k] + hs;
            for (size_t k = 0; k < q.perm.size(); k++)
                m.perm[(size_t)(vs - 4LL * ninst) + k] = (int)pstart[i] + q.perm[k];
            q = mcl::HostMesh();  // release the instance's storage
        }
    };
    pool.clear();
    for (int w = 1; w < nt; w++) pool.emplace_back(copy);
    copy();
    for (auto& th : pool) th.join();
    *out = h;
    return MCL_OK;
}

extern "C" int mcl_hmesh_batch_offsets(const mcl_hmesh_t* h, int* vstart, int* hestart, int* tstart) {
    if (!h || h->m.vstart.empty()) return MCL_ERR_INPUT;
    const size_t n = h->m.vstart.size();
    if (vstart) memcpy(vstart, h->m.vstart.data(), sizeof(int) * n);
    if (hestart) memcpy(hestart, h->m.hestart.data(), sizeof(int) * n);
    if (tstart) memcpy(tstart, h->m.tstart.data(), sizeof(int) * n);
    return MCL_OK;
}

extern "C" int mcl_hmesh_sizes(const mcl_hmesh_t* h, int* npos, int* nv, int* nhe, int* nt) {
    if (!h) return MCL_ERR_INPUT;
    if (npos) *npos = h->m.nv;
    if (nv) *nv = h->m.nv;
    if (nhe) *nhe = (int)(h->m.he.size() / 4);
    if (nt) *nt = (int)h->m.t.size();
    return MCL_OK;
}

extern "C" int mcl_hmesh_export(const mcl_hmesh_t* h, float* pos, int* v2, int* he4, int* t1, int* perm) {
    if (!h) return MCL_ERR_INPUT;
    const mcl::HostMesh& m = h->m;
    if (pos) memcpy(pos, m.pos.data(), sizeof(float) * m.pos.size());
    if (v2) memcpy(v2, m.vrec.data(), sizeof(int) * m.vrec.size());
    if (he4) memcpy(he4, m.he.data(), sizeof(int) * m.he.size());
    if (t1) memcpy(t1, m.t.data(), sizeof(int) * m.t.size());
    if (perm) memcpy(perm, m.perm.data(), sizeof(int) * m.perm.size());
    return MCL_OK;
}

extern "C" int mcl_hmesh_free(mcl_hmesh_t* h) {
    delete h;
    return MCL_OK;
}

extern "C" const char* mcl_version(void) { return "mcleap_hip 0.1 (gfx950)"; }

extern "C" int mcl_device_info(char* name, int len, int* cu_count) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return MCL_ERR_NO_DEVICE;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return MCL_ERR_HIP;
    if (name && len > 0) {
        std::string s = p.gcnArchName;
        strncpy(name, s.c_str(), (size_t)len - 1);
        name[len - 1] = 0;
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    return MCL_OK;
}

// ------------------------------------------------------------ state files
// A triangulation on disk: the AoS arrays of the reference's host object
// (Host_Triangulation.h:27-30, as transferToHost fills them) after a 48-byte
// header, then an FNV-1a 64 checksum of the arrays.
//   char magic[8] "MCLTRI01"; int32 version 1, periodic, npos, nv, nhe, nt,
//   the library's round and stamp counters (0, 0 from mcl_state_save);
//   float pos[2 npos]; int32 v2[2 nv]; int32 he4[4 nhe]; int32 t1[nt];
//   uint64 checksum.  Little-endian.
namespace {
constexpr char kMagic[8] = {'M', 'C', 'L', 'T', 'R', 'I', '0', '1'};
struct StateHeader {
    char magic[8];
    int32_t version, periodic, npos, nv, nhe, nt, reserved[2];
};
static_assert(sizeof(StateHeader) == 40, "header layout");
void fnv(uint64_t& h, const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
}
bool read_header(FILE* f, StateHeader& hd) {
    if (fread(&hd, sizeof(hd), 1, f) != 1) return false;
    if (memcmp(hd.magic, kMagic, 8) != 0 || hd.version != 1) return false;
    if (hd.npos < 0 || hd.nv < 0 || hd.nhe < 0 || hd.nt < 0 || (hd.nhe & 1)) return false;
    return true;
}
}  // namespace

static int state_write(const char* path, int periodic, const float* pos, int npos, const int* v2, int nv,
                       const int* he4, int nhe, const int* t1, int nt, int r0, int r1) {
    if (!path || npos < 0 || nv < 0 || nhe < 0 || nt < 0 || (npos && !pos) || (nv && !v2) || (nhe && !he4) ||
        (nt && !t1))
        return MCL_ERR_INPUT;
    FILE* f = fopen(path, "wb");
    if (!f) return MCL_ERR_INPUT;
    StateHeader hd;
    memcpy(hd.magic, kMagic, 8);
    hd.version = 1;
    hd.periodic = periodic ? 1 : 0;
    hd.npos = npos; hd.nv = nv; hd.nhe = nhe; hd.nt = nt;
    hd.reserved[0] = r0;
    hd.reserved[1] = r1;
    uint64_t h = 14695981039346656037ull;
    const struct { const void* p; size_t n; } parts[4] = {
        {pos, sizeof(float) * 2 * (size_t)npos}, {v2, sizeof(int) * 2 * (size_t)nv},
        {he4, sizeof(int) * 4 * (size_t)nhe}, {t1, sizeof(int) * (size_t)nt}};
    bool ok = fwrite(&hd, sizeof(hd), 1, f) == 1;
    for (const auto& q : parts) {
        if (!ok || q.n == 0) continue;
        ok = fwrite(q.p, 1, q.n, f) == q.n;
        fnv(h, q.p, q.n);
    }
    ok = ok && fwrite(&h, sizeof(h), 1, f) == 1;
    ok = (fclose(f) == 0) && ok;
    return ok ? MCL_OK : MCL_ERR_INPUT;
}

extern "C" int mcl_state_save(const char* path, int periodic, const float* pos, int npos, const int* v2, int nv,
                              const int* he4, int nhe, const int* t1, int nt) {
    return state_write(path, periodic, pos, npos, v2, nv, he4, nhe, t1, nt, 0, 0);
}

extern "C" int mcl_state_info(const char* path, int* periodic, int* npos, int* nv, int* nhe, int* nt) {
    if (!path) return MCL_ERR_INPUT;
    FILE* f = fopen(path, "rb");
    if (!f) return MCL_ERR_INPUT;
    StateHeader hd;
    const bool ok = read_header(f, hd);
    fclose(f);
    if (!ok) return MCL_ERR_INPUT;
    if (periodic) *periodic = hd.periodic;
    if (npos) *npos = hd.npos;
    if (nv) *nv = hd.nv;
    if (nhe) *nhe = hd.nhe;
    if (nt) *nt = hd.nt;
    return MCL_OK;
}

extern "C" int mcl_state_load(const char* path, float* pos, int* v2, int* he4, int* t1) {
    if (!path) return MCL_ERR_INPUT;
    FILE* f = fopen(path, "rb");
    if (!f) return MCL_ERR_INPUT;
    StateHeader hd;
    bool ok = read_header(f, hd);
    uint64_t h = 14695981039346656037ull, stored = 0;
    if (ok) {
        const struct { void* p; size_t n; } parts[4] = {
            {pos, sizeof(float) * 2 * (size_t)hd.npos}, {v2, sizeof(int) * 2 * (size_t)hd.nv},
            {he4, sizeof(int) * 4 * (size_t)hd.nhe}, {t1, sizeof(int) * (size_t)hd.nt}};
        std::vector<unsigned char> tmp;
        for (const auto& q : parts) {
            if (!ok || q.n == 0) continue;
            void* dst = q.p;
            if (!dst) {  // not wanted: read into scratch for the checksum
                tmp.resize(q.n);
                dst = tmp.data();
            }
            ok = fread(dst, 1, q.n, f) == q.n;
            if (ok) fnv(h, dst, q.n);
        }
        ok = ok && fread(&stored, sizeof(stored), 1, f) == 1 && stored == h && fgetc(f) == EOF;
    }
    fclose(f);
    return ok ? MCL_OK : MCL_ERR_INPUT;
}

extern "C" int mcl_tri_save(mcl_tri_t* t, const char* path) {
    if (!t || !path) return MCL_ERR_INPUT;
    int npos = 0, nv = 0, nhe = 0, nt = 0, periodic = 0;
    int rc = mcl_tri_sizes(t, &npos, &nv, &nhe, &nt);
    if (rc != MCL_OK) return rc;
    rc = mcl_tri_is_periodic(t, &periodic);
    if (rc != MCL_OK) return rc;
    std::vector<float> pos(2 * (size_t)npos);
    std::vector<int> v2(2 * (size_t)nv), he4(4 * (size_t)nhe), t1((size_t)nt);
    rc = mcl_tri_download(t, pos.data(), v2.data(), he4.data(), t1.data());
    if (rc != MCL_OK) return rc;
    int round = 0, stamp = 0;
    mcl::tri_counters(t, &round, &stamp, false);
    return state_write(path, periodic, pos.data(), npos, v2.data(), nv, he4.data(), nhe, t1.data(), nt, round, stamp);
}

extern "C" int mcl_tri_load(const char* path, mcl_tri_t** out) {
    if (!path || !out) return MCL_ERR_INPUT;
    int periodic = 0, npos = 0, nv = 0, nhe = 0, nt = 0;
    int rc = mcl_state_info(path, &periodic, &npos, &nv, &nhe, &nt);
    if (rc != MCL_OK) return rc;
    std::vector<float> pos(2 * (size_t)npos);
    std::vector<int> v2(2 * (size_t)nv), he4(4 * (size_t)nhe), t1((size_t)nt);
    rc = mcl_state_load(path, pos.data(), v2.data(), he4.data(), t1.data());
    if (rc != MCL_OK) return rc;
    StateHeader hd;
    FILE* f = fopen(path, "rb");
    const bool ok = f && read_header(f, hd);
    if (f) fclose(f);
    if (!ok) return MCL_ERR_INPUT;
    rc = mcl_tri_create(pos.data(), npos, v2.data(), nv, he4.data(), nhe, t1.data(), nt, periodic, out);
    if (rc != MCL_OK) return rc;
    // the flip rounds' keys and walk seeds continue where the saved run was
    if (hd.reserved[0] > 0 && hd.reserved[1] > 0) mcl::tri_counters(*out, &hd.reserved[0], &hd.reserved[1], true);
    return MCL_OK;
}
