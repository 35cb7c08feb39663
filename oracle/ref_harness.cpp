// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// C entry points over the reference's own kernel-argument producers, compiled
// from their unmodified sources under /root/reference/core by build_ref.sh:
//   TransferFunction::GetTransferFunction (core/TransferFunction.cpp:38-140)
//   RenderInfo + Methods::ComputeTransformMatrix (core/RenderInfo.cpp,
//   core/Methods.cpp:52-147)
//   CPR geometry helpers: Methods::RotatePoint3D, Projection_Point2Plane,
//   Distance_Point2Line (core/Methods.cpp:563-633) over Direction3d /
//   Point3d (core/Direction.h, core/Point.h). Directions are built the way
//   CPRInfo builds them, from two points.
//   MPR plane helpers: the (dirH, dirV) overloads of Distance_Point2Plane and
//   Projection_Point2Plane (core/Methods.cpp:582-598), GetTransferPoint
//   (core/Methods.cpp:667-679), PlaneInfo::GetNormDirection / GetLeftTopPoint
//   (core/PlaneInfo.h:37-45).
//   fpng_encode_image_to_memory (core/fpng/fpng.cpp): the reference's PNG
//   encoder, as the CPU baseline of the device encoder.
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
#include "PlaneInfo.h"
#include "TransferFunction.h"
#include "fpng/fpng.h"

#include <vector>

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

static Point3d P3(const double* p) { return Point3d(p[0], p[1], p[2]); }
static void out3(const Point3d& p, double* o) {
    Point3d q = p;
    o[0] = q.x();
    o[1] = q.y();
    o[2] = q.z();
}

// direction = Direction3d(a, b)
void ref_direction(const double* a, const double* b, double* o) {
    Direction3d d(P3(a), P3(b));
    o[0] = d.x();
    o[1] = d.y();
    o[2] = d.z();
}
// Direction3d(a0, a1).cross(Direction3d(b0, b1))
void ref_cross(const double* a0, const double* a1, const double* b0, const double* b1, double* o) {
    Direction3d d = Direction3d(P3(a0), P3(a1)).cross(Direction3d(P3(b0), P3(b1)));
    o[0] = d.x();
    o[1] = d.y();
    o[2] = d.z();
}
void ref_rotate_point(const double* pt, const double* ax0, const double* ax1, const double* center, float theta,
                      double* o) {
    out3(Methods::RotatePoint3D(P3(pt), Direction3d(P3(ax0), P3(ax1)), P3(center), theta), o);
}
void ref_projection_point_to_plane(const double* pt, const double* n0, const double* n1, const double* pp, double* o) {
    out3(Methods::Projection_Point2Plane(P3(pt), Direction3d(P3(n0), P3(n1)), P3(pp)), o);
}
double ref_distance_point_to_line(const double* pt, const double* d0, const double* d1, const double* pl) {
    return Methods::Distance_Point2Line(P3(pt), Direction3d(P3(d0), P3(d1)), P3(pl));
}
double ref_patient_distance(const double* a, const double* b, const double* sp) {
    return P3(a).PatientDistanceTo(P3(b), P3(sp));
}

double ref_distance_point_to_plane_hv(const double* pt, const double* h0, const double* h1, const double* v0,
                                      const double* v1, const double* pp) {
    return Methods::Distance_Point2Plane(P3(pt), Direction3d(P3(h0), P3(h1)), Direction3d(P3(v0), P3(v1)), P3(pp));
}
void ref_projection_point_to_plane_hv(const double* pt, const double* h0, const double* h1, const double* v0,
                                      const double* v1, const double* pp, double* o) {
    out3(Methods::Projection_Point2Plane(P3(pt), Direction3d(P3(h0), P3(h1)), Direction3d(P3(v0), P3(v1)), P3(pp)), o);
}
void ref_transfer_point(const double* m9, const double* pt, double* o) {
    double m[3][3];
    for (int i = 0; i < 9; ++i) m[i / 3][i % 3] = m9[i];
    out3(Methods::GetTransferPoint(m, P3(pt)), o);
}
// PlaneInfo with dirH = Direction3d(h0, h1), dirV = Direction3d(v0, v1): its
// normal and the left-top corner for a centre point
void ref_plane_info(const double* h0, const double* h1, const double* v0, const double* v1, int w, int h, double ps,
                    const double* center, double* norm, double* left_top) {
    PlaneInfo info;
    info.m_dirH = Direction3d(P3(h0), P3(h1));
    info.m_dirV = Direction3d(P3(v0), P3(v1));
    info.m_nWidth = w;
    info.m_nHeight = h;
    info.m_fPixelSpacing = ps;
    Direction3d n = info.GetNormDirection();
    norm[0] = n.x();
    norm[1] = n.y();
    norm[2] = n.z();
    out3(info.GetLeftTopPoint(P3(center)), left_top);
}

// fpng encode into out (capacity cap); returns the PNG size, 0 on failure or
// when cap is too small.
unsigned long long ref_fpng_encode(const unsigned char* img, int w, int h, int chans, unsigned char* out,
                                   unsigned long long cap) {
    static bool inited = false;
    if (!inited) {
        fpng::fpng_init();
        inited = true;
    }
    std::vector<uint8_t> buf;
    if (!fpng::fpng_encode_image_to_memory(img, (uint32_t)w, (uint32_t)h, (uint32_t)chans, buf)) return 0;
    if (buf.size() > cap) return 0;
    std::memcpy(out, buf.data(), buf.size());
    return buf.size();
}

}  // extern "C"
