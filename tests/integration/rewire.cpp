// The reference-side rewiring of the CUDA layer onto the C ABI, as a
// maintainer would write it in MonkeyGL's own classes:
//   CuDataManager (core/CuDataManager.cpp:34-88): construction, destruction,
//     CopyVolumeData, CopyMaskData, SetTransferFunction;
//   VRImageLayer::GetRGBData's parameter fill and cu_render call
//     (core/VRImageLayer.cpp:77-97) -> mk_render;
//   MPRImageLayer's cu_renderPlane_* calls (core/MPRImageLayer.cpp:144-194)
//     -> mk_render_plane;
//   CPRImageLayer's cu_renderCPR call (core/CPRImageLayer.cpp:67-75)
//     -> mk_render_cpr.
// It is compiled against the reference's own core/Defines.h, so the types
// handed across (RGBA, VOI, AlphaAndWWWLInfo, RenderType, MPRType) are the
// reference's, and against include/monkeygl_hip.h; it links against
// libmonkeygl_hip.so. The classes around the call sites (DataManager,
// RenderInfo, the layers) include cuda_runtime.h in the reference, so minimal
// stand-ins supply the values those call sites read.
//
// Run without a GPU it reports the context creation's status (MK_ERR_HIP)
// and exits 0; with a GPU it renders a small volume through every call site
// and prints a checksum of each image.
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "Defines.h"
#include "monkeygl_hip.h"

using namespace MonkeyGL;

// the reference's structs cross the ABI by memcpy: their layouts must agree
static_assert(sizeof(AlphaAndWWWLInfo) == sizeof(mk_alpha_wwwl) * MK_NUM_LABELS, "AlphaAndWWWLInfo layout");
static_assert(sizeof(VOI) == sizeof(mk_voi), "VOI layout");
static_assert(MAXOBJECTCOUNT == MK_MAX_OBJECTS, "label count");
static_assert((int)RenderTypeVR == MK_RENDER_VR && (int)RenderTypeMIP == MK_RENDER_MIP &&
                  (int)RenderTypeSurface == MK_RENDER_SURFACE, "RenderType values");
static_assert((int)MPRTypeAverage == MK_MPR_AVERAGE && (int)MPRTypeMIP == MK_MPR_MIP &&
                  (int)MPRTypeMinIP == MK_MPR_MINIP, "MPRType values");

struct float3 { float x, y, z; };   // CUDA's vector type, as the reference's DataManager returns it
struct Extent { size_t width, height, depth; };   // cudaExtent's fields

// ---- core/CuDataManager.cpp, rewired -------------------------------------
class CuDataManager {
public:
    CuDataManager() { m_status = mk_context_create(-1, &m_ctx); }   // -1: the device with the most memory
    ~CuDataManager() { mk_context_destroy(m_ctx); }
    bool CopyVolumeData(short* h, Extent s) {                        // was cu_copyVolumeData
        return mk_copy_volume_data(m_ctx, h, (int)s.width, (int)s.height, (int)s.depth) == MK_OK;
    }
    bool CopyMaskData(unsigned char* h, Extent s) {                  // was cu_copyMaskData
        return mk_copy_mask_data(m_ctx, h, (int)s.width, (int)s.height, (int)s.depth) == MK_OK;
    }
    bool SetTransferFunction(float* lut, int n, unsigned char label) {   // was cu_setTransferFunc
        return mk_set_transfer_func(m_ctx, label, lut, n) == MK_OK;
    }
    const float* GetTransformMatrix() const { return m_M; }
    mk_context* context() { return m_ctx; }
    mk_status status() const { return m_status; }

private:
    mk_context* m_ctx = nullptr;
    mk_status m_status = MK_OK;
    float m_M[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
};

// ---- stand-ins for what the layers read ----------------------------------
struct DataManagerStandIn {
    int dims[3] = {48, 40, 36};
    float3 spacing{1.0f, 1.0f, 1.2f};
    RenderType GetRenderType() const { return RenderTypeVR; }
    bool Need2InvertZ() const { return false; }
    float3 GetSpacing() const { return spacing; }
    float3 GetSpacingVoxel() const { return {1.0f / dims[0], 1.0f / dims[1], 1.0f / dims[2]}; }
    float3 GetMaxLenSpacing() const {   // DataManager::InitCommon: extent over the largest extent
        float e[3] = {dims[0] * spacing.x, dims[1] * spacing.y, dims[2] * spacing.z};
        float m = e[0] > e[1] ? (e[0] > e[2] ? e[0] : e[2]) : (e[1] > e[2] ? e[1] : e[2]);
        return {m / e[0], m / e[1], m / e[2]};
    }
    VOI GetVOINormalize() const {
        VOI v;
        v.left = 0; v.right = dims[0] - 1; v.posterior = 0; v.anterior = dims[1] - 1; v.head = 0; v.foot = dims[2] - 1;
        return v;
    }
    RGBA GetColorBackground() const { return RGBA(0.0f, 0.0f, 0.0f, 1.0f); }
    AlphaAndWWWLInfo GetAlphaAndWWWLInfo() const {
        AlphaAndWWWLInfo aw;
        aw.m[0] = AlphaAndWWWL(1.0f, 2000.0f, 0.0f);
        return aw;
    }
};
struct RenderInfoStandIn {
    float GetTotalXTranslate() const { return 0.0f; }
    float GetTotalYTranslate() const { return 0.0f; }
    float GetTotalScale() const { return 1.0f; }
};

// ---- core/VRImageLayer.cpp:77-97, rewired ----------------------------------
static bool vr_layer(CuDataManager& cu, const DataManagerStandIn* dm, const RenderInfoStandIn& ri, int nWidth,
                     int nHeight, std::unique_ptr<unsigned char[]>& pData) {
    pData.reset(new unsigned char[(size_t)nWidth * nHeight * 3]);
    mk_render_params p = {};
    p.width = nWidth;  p.height = nHeight;
    p.render_type = (int)dm->GetRenderType();
    p.invert_z = dm->Need2InvertZ();
    float3 mls = dm->GetMaxLenSpacing(), sp = dm->GetSpacing(), sv = dm->GetSpacingVoxel();
    memcpy(p.max_len_spacing, &mls, 12); memcpy(p.spacing, &sp, 12); memcpy(p.spacing_voxel, &sv, 12);
    p.x_translate = ri.GetTotalXTranslate(); p.y_translate = ri.GetTotalYTranslate(); p.scale = ri.GetTotalScale();
    memcpy(p.transform, cu.GetTransformMatrix(), 36);
    VOI voi = dm->GetVOINormalize(); memcpy(&p.voi, &voi, sizeof(mk_voi));
    RGBA bg = dm->GetColorBackground();
    p.background[0] = bg.red; p.background[1] = bg.green; p.background[2] = bg.blue; p.background[3] = bg.alpha;
    AlphaAndWWWLInfo aw = dm->GetAlphaAndWWWLInfo(); memcpy(p.alpha_wwwl, aw.m, sizeof(p.alpha_wwwl));
    p.step_scale = 1.0f; p.exit_alpha = 0.995f;
    return mk_render(cu.context(), &p, pData.get()) == MK_OK;   // was cu_render(...)
}

// ---- core/MPRImageLayer.cpp:144-194, rewired (axial plane) ------------------
static bool mpr_layer(CuDataManager& cu, const DataManagerStandIn* dm, MPRType type, float halfNum,
                      std::unique_ptr<short[]>& pData, int& w, int& h) {
    w = dm->dims[0]; h = dm->dims[1];
    pData.reset(new short[(size_t)w * h]);
    float3 sp = dm->GetSpacing();
    mk_plane_params q = {};
    q.width = w; q.height = h;
    q.mpr_type = (int)type;
    q.sampling = MK_SAMPLE_LINEAR;
    q.invert_z = dm->Need2InvertZ();
    memcpy(q.spacing, &sp, 12);
    const float dirH[3] = {1, 0, 0}, dirV[3] = {0, 1, 0}, dirN[3] = {0, 0, 1};
    memcpy(q.dir_h, dirH, 12); memcpy(q.dir_v, dirV, 12); memcpy(q.dir_n, dirN, 12);
    const float lt[3] = {0.0f, 0.0f, 0.5f * dm->dims[2] * sp.z};
    memcpy(q.pt_left_top, lt, 12);
    q.pixel_spacing = sp.x; q.half_num = halfNum;
    return mk_render_plane(cu.context(), &q, pData.get()) == MK_OK;   // replaces cu_renderPlane_*
}

// ---- core/CPRImageLayer.cpp:67-75, rewired ---------------------------------
static bool cpr_layer(CuDataManager& cu, const DataManagerStandIn* dm, int nWidth, int nHeight,
                      std::unique_ptr<short[]>& pData) {
    pData.reset(new short[(size_t)nWidth * nHeight]);
    std::vector<double> pPoints((size_t)nHeight * 3), pDirs((size_t)nHeight * 3);
    for (int i = 0; i < nHeight; ++i) {   // a straight centreline along z, sampled across x
        pPoints[3 * i + 0] = 0.0; pPoints[3 * i + 1] = 0.5 * dm->dims[1]; pPoints[3 * i + 2] = i + 0.5;
        pDirs[3 * i + 0] = 1.0; pDirs[3 * i + 1] = 0.0; pDirs[3 * i + 2] = 0.0;
    }
    return mk_render_cpr(cu.context(), nWidth, nHeight, pPoints.data(), pDirs.data(), dm->Need2InvertZ(),
                         MK_SAMPLE_LINEAR, pData.get()) == MK_OK;   // was cu_renderCPR
}

int main() {
    CuDataManager cu;
    if (cu.status() != MK_OK) {
        printf("context: status %d: %s\n", (int)cu.status(), mk_last_error(nullptr));
        // without a GPU the library reports MK_ERR_HIP and the process goes on
        return (cu.status() == MK_ERR_HIP && mk_device_count() == 0) ? 0 : 1;
    }
    DataManagerStandIn dm;
    const size_t n = (size_t)dm.dims[0] * dm.dims[1] * dm.dims[2];
    std::vector<short> vol(n);
    for (size_t i = 0; i < n; ++i) vol[i] = (short)((i * 2654435761u >> 7) % 2000) - 1000;
    std::vector<float> lut(100 * 4);
    for (int i = 0; i < 100; ++i) { lut[4 * i] = lut[4 * i + 1] = lut[4 * i + 2] = 1.0f; lut[4 * i + 3] = i / 400.0f; }
    Extent e{(size_t)dm.dims[0], (size_t)dm.dims[1], (size_t)dm.dims[2]};
    if (!cu.CopyVolumeData(vol.data(), e) || !cu.SetTransferFunction(lut.data(), 100, 0)) {
        printf("upload failed: %s\n", mk_last_error(cu.context()));
        return 1;
    }
    RenderInfoStandIn ri;
    std::unique_ptr<unsigned char[]> rgb;
    std::unique_ptr<short[]> mpr, cpr;
    int w = 0, h = 0;
    if (!vr_layer(cu, &dm, ri, 64, 64, rgb) || !mpr_layer(cu, &dm, MPRTypeMIP, 2.0f, mpr, w, h) ||
        !cpr_layer(cu, &dm, 48, 36, cpr)) {
        printf("render failed: %s\n", mk_last_error(cu.context()));
        return 1;
    }
    unsigned long long s0 = 0, s1 = 0, s2 = 0;
    for (int i = 0; i < 64 * 64 * 3; ++i) s0 += rgb[i];
    for (int i = 0; i < w * h; ++i) s1 += (unsigned short)mpr[i];
    for (int i = 0; i < 48 * 36; ++i) s2 += (unsigned short)cpr[i];
    printf("vr %llu mpr %llu cpr %llu\n", s0, s1, s2);
    return 0;
}
