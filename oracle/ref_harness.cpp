// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// C entry points over the reference's own kernel-argument producers, compiled
// from their unmodified sources under /root/reference/core by build_ref.sh:
//   TransferFunction::GetTransferFunction (core/TransferFunction.cpp:38-140)
//   RenderInfo + Methods::ComputeTransformMatrix (core/RenderInfo.cpp,
//   core/Methods.cpp:52-147)
// Used only to write golden vectors (tests/golden/make_golden.py) and to
// cross-check the restatements where the reference is present.
#include <cstring>
#include <map>
#include <memory>

// RenderInfo keeps its matrices private (DataManager is its friend); the
// harness reads them through this access override, which changes no layout.
#define private public
#include "RenderInfo.h"
#undef private
#include "Methods.h"
#include "TransferFunction.h"

using namespace MonkeyGL;

extern "C" {

int ref_transfer_function(const int* pos, const float* rgba, int n, const int* apos, const float* alpha, int na,
                          float* out, int* out_len) {
    std::map<int, RGBA> pts;
    for (int i = 0; i < n; ++i) pts[pos[i]] = RGBA(rgba[4 * i], rgba[4 * i + 1], rgba[4 * i + 2], rgba[4 * i + 3]);
    std::map<int, float> ap;
    for (int i = 0; i < na; ++i) ap[apos[i]] = alpha[i];
    TransferFunction tf;
    tf.SetControlPoints(pts, ap);
    std::shared_ptr<RGBA> buf;
    int len = 0;
    if (!tf.GetTransferFunction(buf, len)) return 0;
    std::memcpy(out, buf.get(), sizeof(RGBA) * len);
    *out_len = len;
    return 1;
}

void* ref_render_info_new() { return new RenderInfo(); }
void ref_render_info_free(void* h) { delete (RenderInfo*)h; }
void ref_render_info_rotate(void* h, float x, float y) { ((RenderInfo*)h)->Rotate(x, y); }
float ref_render_info_zoom(void* h, float r) { return ((RenderInfo*)h)->Zoom(r); }
void ref_render_info_pan(void* h, float x, float y) { ((RenderInfo*)h)->Pan(x, y); }
void ref_render_info_preset(void* h, int which) {
    RenderInfo* r = (RenderInfo*)h;
    switch (which) {
        case 0: r->Anterior(); break;
        case 1: r->Posterior(); break;
        case 2: r->Left(); break;
        case 3: r->Right(); break;
        case 4: r->Head(); break;
        default: r->Foot(); break;
    }
}
void ref_render_info_set_vr_size(void* h, int w, int hh, int* ow, int* oh) {
    RenderInfo* r = (RenderInfo*)h;
    r->SetVRSize(w, hh);
    r->GetVRSize(*ow, *oh);
}
// m, m_t, rot, rot_t: float[9] each; xyz: translate x, y and total scale
void ref_render_info_get(void* h, float* m, float* m_t, float* rot, float* rot_t, float* xyz) {
    RenderInfo* r = (RenderInfo*)h;
    std::memcpy(m, r->m_pTransformMatrix, 9 * sizeof(float));
    std::memcpy(m_t, r->m_pTransposeTransformMatrix, 9 * sizeof(float));
    std::memcpy(rot, r->m_pRotateMatrix, 9 * sizeof(float));
    std::memcpy(rot_t, r->m_pTransposRotateMatrix, 9 * sizeof(float));
    xyz[0] = r->GetTotalXTranslate();
    xyz[1] = r->GetTotalYTranslate();
    xyz[2] = r->GetTotalScale();
}

}  // extern "C"
