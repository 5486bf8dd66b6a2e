// mcl_api.hip -- host-mesh construction entry points and library info.
#include <cstring>
#include <string>

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
