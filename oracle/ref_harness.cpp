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
//   The orientation figure of VR images: Methods::FillHoleInImage_Ch1
//   (core/Methods.cpp:308-368) under the few lines of
//   VRImageLayer::MergeOrientationBox / GetMeshPoints and
//   MarchingCube::LoadBinarySTL that call it (restated below: their files
//   do not build here); the CPR line of VR images (TransferVoxel2ImageInVR
//   over Methods::MatrixMul, DrawLineInImage24Bit).
// Used only to write golden vectors (tests/golden/make_golden.py) and to
// cross-check the restatements where the reference is present.
#include <cfloat>
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

// MarchingCube::LoadBinarySTL's normalisation (core/MarchingCube.cpp:306-365;
// MarchingCube.cpp itself links the reference's logger, absent here) over the
// reference's Facet3D / Point3f: body is the STL after its 84-byte header.
void ref_stl_normalize(const unsigned char* body, int n, float* out9) {
    std::vector<Facet3D> facets;
    float xmin = FLT_MAX, xmax = FLT_MIN;
    float ymin = FLT_MAX, ymax = FLT_MIN;
    float zmin = FLT_MAX, zmax = FLT_MIN;
    for (int i = 0; i < n; i++) {
        Facet3D facet;
        memcpy(&facet, body + (size_t)i * 50, 50);
        facets.push_back(facet);
        xmin = min(xmin, min(facet.v1.x(), min(facet.v2.x(), facet.v3.x())));
        xmax = max(xmax, max(facet.v1.x(), max(facet.v2.x(), facet.v3.x())));
        ymin = min(ymin, min(facet.v1.y(), min(facet.v2.y(), facet.v3.y())));
        ymax = max(ymax, max(facet.v1.y(), max(facet.v2.y(), facet.v3.y())));
        zmin = min(zmin, min(facet.v1.z(), min(facet.v2.z(), facet.v3.z())));
        zmax = max(zmax, max(facet.v1.z(), max(facet.v2.z(), facet.v3.z())));
    }
    float xlen = xmax - xmin;
    float zlen = zmax - zmin;
    float maxlen = max(xlen, zlen);
    Point3f ptCenter((xmin + xmax) / 2, (ymin + ymax) / 2, (zmin + zmax) / 2);
    for (size_t i = 0; i < facets.size(); i++) {
        Facet3D& facet = facets[i];
        Point3f* v[3] = {&facet.v1, &facet.v2, &facet.v3};
        for (int k = 0; k < 3; ++k) {
            *v[k] -= ptCenter;
            *v[k] /= maxlen;
            v[k]->SetZ(-v[k]->z());
            out9[i * 9 + k * 3 + 0] = v[k]->x();
            out9[i * 9 + k * 3 + 1] = v[k]->y();
            out9[i * 9 + k * 3 + 2] = v[k]->z();
        }
    }
}

// VRImageLayer::MergeOrientationBox + GetMeshPoints + the file's
// TransferPoint2D (core/VRImageLayer.cpp:104-194; VRImageLayer.cpp pulls in
// the CUDA data manager, so its few lines are restated here) over the
// reference's Methods::GetTransferPointf, Direction3f, Point2f, RGBA and
// Methods::FillHoleInImage_Ch1 (core/Methods.cpp:308-368), compiled from its
// source. facets9: normalised facets; rgb: the VR image, modified in place.
static Point2f ref_transfer_point_2d(Point2f pt, int nWidth, int nHeight) {
    float r = nHeight;
    Point2f ptOut = pt;
    ptOut *= r;
    ptOut += Point2f(nWidth / 2.0, nHeight / 2.0);
    return ptOut;
}

void ref_orientation_box(const float* facets9, int n, const float* rot9, unsigned char* pVR, int nWidth, int nHeight) {
    int nW_Box = 64, nH_Box = 64;
    if (nWidth <= nW_Box || nHeight <= nH_Box) return;
    int w = 512, h = 512;
    float rot[9];
    memcpy(rot, rot9, sizeof(rot));
    std::vector<float> image(w * h, 0.0f), zb(w * h, 0.0f);
    for (int i = 0; i < n; i++) {
        const float* q = facets9 + (size_t)i * 9;
        Point3f v1 = Methods::GetTransferPointf(rot, Point3f(q[0], q[1], q[2]));
        Point3f v2 = Methods::GetTransferPointf(rot, Point3f(q[3], q[4], q[5]));
        Point3f v3 = Methods::GetTransferPointf(rot, Point3f(q[6], q[7], q[8]));
        Direction3f nrm = Direction3f(v3, v1).cross(Direction3f(v3, v2));
        float d = nrm.dot(Direction3f(0, -1, 0));
        if (d < 0) continue;
        float zBuffer = (v1.y() + v2.y() + v3.y()) / 3 + 10000;
        Point2f p1 = ref_transfer_point_2d(Point2f(v1.x(), v1.z()), w, h);
        Point2f p2 = ref_transfer_point_2d(Point2f(v2.x(), v2.z()), w, h);
        Point2f p3 = ref_transfer_point_2d(Point2f(v3.x(), v3.z()), w, h);
        Methods::FillHoleInImage_Ch1(image.data(), zb.data(), w, h, d, zBuffer, p1, p2, p3);
    }
    RGBA clr(0.902, 0.902, 0.302);
    for (int y = nHeight - nH_Box; y < nHeight; y++) {
        int yIdx = (y + nH_Box - nHeight) * 8;
        for (int x = nWidth - nW_Box; x < nWidth; x++) {
            int xIdx = (x + nW_Box - nWidth) * 8;
            float diffuse = image[yIdx * w + xIdx];
            if (diffuse > 0) {
                RGBA clrTemp = clr * diffuse;
                clrTemp = clrTemp + 0.05;
                pVR[3 * (y * nWidth + x)] = int(clrTemp.red * 255);
                pVR[3 * (y * nWidth + x) + 1] = int(clrTemp.green * 255);
                pVR[3 * (y * nWidth + x) + 2] = int(clrTemp.blue * 255);
            }
        }
    }
}

// DataManager::TransferVoxel2ImageInVR (core/DataManager.cpp:236-258; the
// file does not build here) over the reference's Point3d and
// Methods::MatrixMul, then CPRLineLayer::GetRGBData's polyline
// (core/CPRLineLayer.cpp:40-65) through Methods::DrawLineInImage24Bit.
// mt9: RenderInfo's transposed transform matrix.
void ref_voxel_to_vr(const double* vox, int n, const int* dims, const double* spacing, int invert_z, const float* mt9,
                     float xt, float yt, int nWidth, int nHeight, float* xy) {
    float mt[9];
    memcpy(mt, mt9, sizeof(mt));
    for (int i = 0; i < n; ++i) {
        double fxSpacing = spacing[0], fySpacing = spacing[1], fzSpacing = spacing[2];
        float fMaxLen = max(dims[0] * fxSpacing, max(dims[1] * fySpacing, dims[2] * fzSpacing));
        Point3d maxper(fMaxLen / (dims[0] * fxSpacing), fMaxLen / (dims[1] * fySpacing),
                       fMaxLen / (dims[2] * fzSpacing));
        Point3d ptVoxel(vox[3 * i], vox[3 * i + 1], vox[3 * i + 2]);
        Point3d pt(ptVoxel.x() / dims[0], ptVoxel.y() / dims[1], ptVoxel.z() / dims[2]);
        if (invert_z) pt.SetZ(1.0 - pt.z());
        pt -= Point3d(0.5, 0.5, 0.5);
        pt /= maxper;
        pt = Methods::MatrixMul(mt, pt);
        xy[2 * i] = (pt.x() + 0.5) * nWidth + xt;
        xy[2 * i + 1] = (pt.z() + 0.5) * nHeight + yt;
    }
}

void ref_polyline_24bit(const float* xy, int n, const float* rgb, unsigned char* img, int w, int h) {
    RGBA clr(rgb[0], rgb[1], rgb[2]);
    if (n < 2) return;
    float x0 = xy[0], y0 = xy[1];
    for (int i = 1; i < n; i++) {
        float x1 = xy[2 * i], y1 = xy[2 * i + 1];
        Methods::DrawLineInImage24Bit(img, w, h, x0, y0, x1, y1, 1, clr);
        x0 = x1;
        y0 = y1;
    }
}

}  // extern "C"
