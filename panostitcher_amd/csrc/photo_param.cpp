// photo_param.cpp -- rig files of the stitch path: the camera parameters that
// prepare() turns into maps.
//
//   XML   loadPhotoParamFromXML (Warp/PhotoParam.cpp:506-511) tries the
//         per-camera layout (:273-504: Root/POSITION|VIDEO, every field in each
//         camera element) and then the shared-lens layout (:17-271: one or more
//         Root elements, each with GlobalInfo holding the lens and optionally the
//         crop, then POSITION|VIDEO elements with the crop if GlobalInfo has none
//         and yaw/pitch/roll).  A missing element rejects the layout; an element
//         whose text is absent or not a number is an error of the whole load
//         (ticpp's GetText throws there, :59 and the like, outside the try).
//   PTS   loadPhotoParamFromPTS (:681-754): after the "ptGui project file" line,
//         the first 'o' line holding all of v a b c d e g t gives the shared
//         lens (and is not a camera, :709-716); every later 'o' line holding
//         f y p r is a camera (projection f, yaw, pitch, roll, optional crop
//         C<left>,<right>,<top>,<bottom>).  The reference's scanning is kept as
//         it is: each key is the first occurrence of its letter anywhere in the
//         line, and a value runs to the end of the line unless the line's first
//         space comes after the key (getValueString, :518-525).
//   dispatch  loadPhotoParams (:756-770): extension "pts" -> PTS, else XML.
//   export    exportPhotoParamToXML (:772-925): the per-camera layout, numbers
//         printed the way an ostream prints them (6 significant digits).
//   rotate    rotateCamera (:931-960): R(yaw, pitch, roll) * R(camera) with
//         setRotationRM / getRotationRM (Warp/Rotation.h:88-111); the rotation
//         is given in radians, camera angles are degrees.
//
// The XML reader below is this library's own: elements, attributes (skipped),
// text with whitespace condensed, CDATA, comments, declarations; no DTDs.
// ticpp/TinyXML is not part of this build.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/panohip.h"

namespace {

using Param = pano_photo_param;

// ------------------------------------------------------------------ XML ----

struct XNode {
    bool text = false;       // a text (or CDATA) node
    std::string value;       // element name or text
    std::vector<XNode> kids;
};

struct XmlReader {
    static constexpr int kMaxDepth = 64;   // deeper nesting is refused (no rig file nests so deep)
    const std::string& s;
    size_t p = 0;
    bool ok = true;
    int depth = 0;

    explicit XmlReader(const std::string& src) : s(src) {}

    bool at(const char* t) const { return s.compare(p, std::strlen(t), t) == 0; }
    bool space(char c) const { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }
    void skip_space()
    {
        while (p < s.size() && space(s[p])) p++;
    }
    bool skip_past(const char* t)
    {
        const size_t q = s.find(t, p);
        if (q == std::string::npos) return ok = false;
        p = q + std::strlen(t);
        return true;
    }
    static bool name_char(char c)
    {
        return std::isalnum((unsigned char)c) || c == '_' || c == '-' || c == '.' || c == ':' ||
               (unsigned char)c >= 0x80;
    }
    // Text up to the next '<': entities decoded, whitespace runs made one
    // space, leading and trailing whitespace dropped.
    std::string read_text()
    {
        std::string out;
        bool ws = false;
        while (p < s.size() && s[p] != '<') {
            char c = s[p];
            if (space(c)) {
                ws = !out.empty();
                p++;
                continue;
            }
            if (ws) out += ' ';
            ws = false;
            if (c == '&') {
                static const struct { const char* e; char c; } ents[] = {
                    {"&amp;", '&'}, {"&lt;", '<'}, {"&gt;", '>'}, {"&quot;", '"'}, {"&apos;", '\''}};
                bool hit = false;
                for (const auto& e : ents)
                    if (at(e.e)) {
                        out += e.c;
                        p += std::strlen(e.e);
                        hit = true;
                        break;
                    }
                if (!hit) {
                    out += c;
                    p++;
                }
                continue;
            }
            out += c;
            p++;
        }
        return out;
    }
    // Children of an element (or of the document when name is empty) up to
    // its end tag.
    void read_children(XNode& parent, const std::string& name)
    {
        for (;;) {
            skip_space();
            if (p >= s.size()) {
                if (!name.empty()) ok = false;   // unterminated element
                return;
            }
            if (s[p] != '<') {
                if (name.empty()) {   // text outside the root: not XML
                    ok = false;
                    return;
                }
                XNode t;
                t.text = true;
                t.value = read_text();
                if (!t.value.empty()) parent.kids.push_back(std::move(t));
                continue;
            }
            if (at("<?")) {
                if (!skip_past("?>")) return;
            } else if (at("<!--")) {
                if (!skip_past("-->")) return;
            } else if (at("<![CDATA[")) {
                p += 9;
                const size_t q = s.find("]]>", p);
                if (q == std::string::npos || name.empty()) {
                    ok = false;
                    return;
                }
                XNode t;
                t.text = true;
                t.value = s.substr(p, q - p);
                parent.kids.push_back(std::move(t));
                p = q + 3;
            } else if (at("<!")) {
                if (!skip_past(">")) return;
            } else if (at("</")) {
                p += 2;
                const size_t b = p;
                while (p < s.size() && name_char(s[p])) p++;
                const std::string end = s.substr(b, p - b);
                skip_space();
                if (name.empty() || end != name || p >= s.size() || s[p] != '>') {
                    ok = false;
                    return;
                }
                p++;
                return;
            } else {
                XNode e;
                if (!read_element(e)) return;
                parent.kids.push_back(std::move(e));
            }
            if (!ok) return;
        }
    }
    bool read_element(XNode& e)
    {
        if (++depth > kMaxDepth) return ok = false;
        struct Leave {
            int& d;
            ~Leave() { d--; }
        } leave{depth};
        p++;   // '<'
        const size_t b = p;
        while (p < s.size() && name_char(s[p])) p++;
        if (p == b) return ok = false;
        e.value = s.substr(b, p - b);
        // attributes: name = "value" | 'value'
        for (;;) {
            skip_space();
            if (p >= s.size()) return ok = false;
            if (at("/>")) {
                p += 2;
                return true;
            }
            if (s[p] == '>') {
                p++;
                read_children(e, e.value);
                return ok;
            }
            const size_t a = p;
            while (p < s.size() && name_char(s[p])) p++;
            if (p == a) return ok = false;
            skip_space();
            if (p >= s.size() || s[p] != '=') return ok = false;
            p++;
            skip_space();
            if (p >= s.size() || (s[p] != '"' && s[p] != '\'')) return ok = false;
            const char q = s[p++];
            const size_t end = s.find(q, p);
            if (end == std::string::npos) return ok = false;
            p = end + 1;
        }
    }
};

bool load_xml(const char* path, XNode& doc)
{
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    std::string src = ss.str();
    if (src.compare(0, 3, "\xEF\xBB\xBF") == 0) src.erase(0, 3);   // UTF-8 byte order mark
    XmlReader r(src);
    r.read_children(doc, "");
    if (!r.ok) return false;
    int roots = 0;
    for (const XNode& k : doc.kids) roots += !k.text;
    return roots > 0;
}

// i-th and later siblings of kids[i] named like it
const XNode* first_elem(const XNode& parent, const char* name, size_t* index = nullptr)
{
    for (size_t i = 0; i < parent.kids.size(); i++)
        if (!parent.kids[i].text && parent.kids[i].value == name) {
            if (index) *index = i;
            return &parent.kids[i];
        }
    return nullptr;
}

struct FormatError {};

// GetText(&double): the element's first child must be text holding a number
// as an istream reads one; anything else is an error of the whole load.
double number_of(const XNode& e)
{
    if (e.kids.empty() || !e.kids[0].text) throw FormatError{};
    std::istringstream in(e.kids[0].value);
    double v = 0;
    in >> v;
    if (in.fail()) throw FormatError{};
    return v;
}

// Reads the named child of e into v; false when there is no such child.
bool field(const XNode& e, const char* name, double& v)
{
    const XNode* c = first_elem(e, name);
    if (!c) return false;
    v = number_of(*c);
    return true;
}

// The POSITION elements of root, or its VIDEO elements when it has no POSITION.
std::vector<const XNode*> camera_elems(const XNode& root)
{
    std::vector<const XNode*> out;
    const char* name = first_elem(root, "POSITION") ? "POSITION" : (first_elem(root, "VIDEO") ? "VIDEO" : nullptr);
    if (!name) return out;
    for (const XNode& k : root.kids)
        if (!k.text && k.value == name) out.push_back(&k);
    return out;
}

// Lens fields in the order both layouts read them.
bool lens_fields(const XNode& e, Param& q, bool with_type)
{
    double v;
    if (with_type) {
        if (!field(e, "IMAGETYPE", v)) return false;
        q.imageType = (int)v;
        if (!field(e, "CROPMODE", v)) return false;
        q.cropMode = (int)v;
        if (!field(e, "HFOV", v)) return false;
        q.hfov = v;
        if (!field(e, "VFOV", v)) return false;
        q.vfov = v;
    }
    if (!field(e, "Alpha", q.alpha) || !field(e, "Beta", q.beta) || !field(e, "Gamma", q.gamma) ||
        !field(e, "ShiftX", q.shiftX) || !field(e, "ShiftY", q.shiftY) || !field(e, "ShearX", q.shearX) ||
        !field(e, "ShearY", q.shearY))
        return false;
    return true;
}

// CROPLEFT/RIGHT/TOP/BOTTOM -> crop rectangle; right/bottom are exclusive
// edges, widths truncated from the double difference.
bool crop_fields(const XNode& e, Param& q)
{
    double v;
    if (!field(e, "CROPLEFT", v)) return false;
    q.cropX = (int)v;
    if (!field(e, "CROPRIGHT", v)) return false;
    q.cropWidth = (int)(v - q.cropX);
    if (!field(e, "CROPTOP", v)) return false;
    q.cropY = (int)v;
    if (!field(e, "CROPBOTTOM", v)) return false;
    q.cropHeight = (int)(v - q.cropY);
    return true;
}

bool pose_fields(const XNode& e, Param& q)
{
    return field(e, "YAW", q.yaw) && field(e, "PITCH", q.pitch) && field(e, "ROLL", q.roll);
}

// Per-camera layout (PhotoParam.cpp:273-504).
bool xml_per_camera(const XNode& doc, std::vector<Param>& out)
{
    out.clear();
    const XNode* root = first_elem(doc, "Root");
    if (!root) return false;
    const std::vector<const XNode*> cams = camera_elems(*root);
    if (cams.empty()) return false;
    for (const XNode* c : cams) {
        Param q;
        std::memset(&q, 0, sizeof q);
        double v;
        if (!lens_fields(*c, q, true) || !crop_fields(*c, q)) return out.clear(), false;
        if (!field(*c, "CircleX", v)) return out.clear(), false;
        q.circleX = (int)v;
        if (!field(*c, "CircleY", v)) return out.clear(), false;
        q.circleY = (int)v;
        if (!field(*c, "CircleR", v)) return out.clear(), false;
        q.circleR = (int)v;
        if (!pose_fields(*c, q)) return out.clear(), false;
        out.push_back(q);
    }
    return true;
}

// Shared-lens layout (PhotoParam.cpp:17-271): every Root element in turn; the
// crop read from a camera element carries over to the next one of its Root.
bool xml_shared_lens(const XNode& doc, std::vector<Param>& out)
{
    out.clear();
    if (!first_elem(doc, "Root")) return false;
    for (const XNode& root : doc.kids) {
        if (root.text || root.value != "Root") continue;
        Param q;
        std::memset(&q, 0, sizeof q);
        const XNode* g = first_elem(root, "GlobalInfo");
        if (!g) return out.clear(), false;
        double v;
        if (!field(*g, "IMAGETYPE", v)) return out.clear(), false;
        q.imageType = (int)v;
        if (!field(*g, "CROPMODE", v)) return out.clear(), false;
        q.cropMode = (int)v;
        if (!field(*g, "HFOV", q.hfov) || !field(*g, "VFOV", q.vfov)) return out.clear(), false;
        const bool crop_in_root = first_elem(*g, "CROPLEFT") != nullptr;
        if (crop_in_root && !crop_fields(*g, q)) return out.clear(), false;
        if (!lens_fields(*g, q, false)) return out.clear(), false;
        const std::vector<const XNode*> cams = camera_elems(root);
        if (cams.empty()) return out.clear(), false;
        for (const XNode* c : cams) {
            if (!crop_in_root && !crop_fields(*c, q)) return out.clear(), false;
            if (!pose_fields(*c, q)) return out.clear(), false;
            out.push_back(q);
        }
    }
    return true;
}

// 0 on success (out may be empty), PANO_ERR_ARG-like negative on a format error.
int load_xml_params(const char* path, std::vector<Param>& out)
{
    out.clear();
    XNode doc;
    if (!load_xml(path, doc)) return 0;
    try {
        if (xml_per_camera(doc, out)) return 0;
        xml_shared_lens(doc, out);
    } catch (const FormatError&) {
        out.clear();
        return PANO_ERR_FORMAT;
    }
    return 0;
}

// ------------------------------------------------------------------ PTS ----

// A number as `stringstream >> double` reads it from str; 0 when none.
double to_double(const std::string& str)
{
    std::stringstream ss;
    ss << str;
    double v = 0;
    ss >> v;
    return v;
}

// The value text of a key at beg: up to the line's first space when that
// comes after beg, else the rest of the line.
std::string value_string(const std::string& line, size_t beg)
{
    const size_t sp = line.find(' ');
    if (sp == std::string::npos || sp < beg) return line.substr(beg);
    return line.substr(beg, sp - beg);
}

double value_at(const std::string& line, size_t beg) { return to_double(value_string(line, beg)); }

bool pts_global(const std::string& line, Param& g)
{
    if (line.empty() || line[0] != 'o') return false;
    const char keys[] = "vabcdegt";
    size_t pos[8];
    for (int k = 0; k < 8; k++)
        if ((pos[k] = line.find(keys[k])) == std::string::npos) return false;
    std::memset(&g, 0, sizeof g);
    g.hfov = value_at(line, pos[0] + 1);
    g.alpha = value_at(line, pos[1] + 1);
    g.beta = value_at(line, pos[2] + 1);
    g.gamma = value_at(line, pos[3] + 1);
    g.shiftX = value_at(line, pos[4] + 1);
    g.shiftY = value_at(line, pos[5] + 1);
    g.shearX = value_at(line, pos[6] + 1);
    g.shearY = value_at(line, pos[7] + 1);
    return true;
}

bool pts_camera(const std::string& line, Param& q)
{
    if (line.empty() || line[0] != 'o') return false;
    const size_t pf = line.find('f'), py = line.find('y'), pp = line.find('p'), pr = line.find('r'),
                 pc = line.find('C');
    if (pf == std::string::npos || py == std::string::npos || pp == std::string::npos ||
        pr == std::string::npos)
        return false;
    std::memset(&q, 0, sizeof q);
    // panotools projection -> PhotoParam::ImageType: 0 rectilinear, 2 circular
    // fisheye -> 3, 3 full-frame fisheye -> 1; others stay 0
    const int f = (int)(value_at(line, pf + 1) + 0.5);
    q.imageType = f == 2 ? 3 : (f == 3 ? 1 : 0);
    q.yaw = value_at(line, py + 1);
    q.pitch = value_at(line, pp + 1);
    q.roll = value_at(line, pr + 1);
    if (pc != std::string::npos) {
        // the value text split at every comma: exactly four numbers
        const std::string vs = value_string(line, pc + 1);
        std::vector<double> v;
        size_t b = 0;
        for (;;) {
            const size_t c = vs.find(',', b);
            v.push_back(to_double(vs.substr(b, c == std::string::npos ? std::string::npos : c - b)));
            if (c == std::string::npos) break;
            b = c + 1;
        }
        if (v.size() != 4) return false;
        q.cropX = (int)v[0];
        q.cropY = (int)v[2];
        q.cropWidth = (int)(v[1] - v[0]);
        q.cropHeight = (int)(v[3] - v[2]);
    }
    return true;
}

void load_pts_params(const char* path, std::vector<Param>& out)
{
    out.clear();
    std::ifstream f(path);
    if (!f.is_open()) return;
    std::string line;
    bool valid = false;
    while (std::getline(f, line))
        if (line.find("ptGui project file") != std::string::npos) {
            valid = true;
            break;
        }
    if (!valid) return;
    Param g;
    bool have_global = false;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        if (!have_global) {
            have_global = pts_global(line, g);
            if (have_global) continue;   // the lens line is not a camera
        }
        Param q;
        if (pts_camera(line, q)) out.push_back(q);
    }
    if (!have_global) {
        out.clear();
        return;
    }
    for (Param& q : out) {
        q.hfov = g.hfov;
        q.alpha = g.alpha;
        q.beta = g.beta;
        q.gamma = g.gamma;
        q.shiftX = g.shiftX;
        q.shiftY = g.shiftY;
        q.shearX = g.shearX;
        q.shearY = g.shearY;
    }
}

int hand_out(const std::vector<Param>& v, pano_photo_param* out, int cap)
{
    const int n = (int)v.size();
    for (int i = 0; i < n && i < cap; i++) out[i] = v[i];
    return n;
}

// --------------------------------------------------------------- rotate ----

// Rotation.h:88-111 (x left, y down, z toward the viewer; R = Ry Rp Rr)
constexpr double kPi = 3.14159265358979;                      // MathConstant.h:3
constexpr double kDegPerRad = 180.0 / 3.1415926535898;        // PhotoParam.cpp:928
constexpr double kRadPerDeg = 3.1415926535898 / 180.0;        // PhotoParam.cpp:929

void rot_rm(double m[3][3], double yaw, double pitch, double roll)
{
    const double cy = std::cos(yaw), sy = std::sin(yaw), cp = std::cos(pitch), sp = std::sin(pitch),
                 cr = std::cos(roll), sr = std::sin(roll);
    m[0][0] = cy * cr + sy * sp * sr;
    m[0][1] = sy * sp * cr - cy * sr;
    m[0][2] = sy * cp;
    m[1][0] = cp * sr;
    m[1][1] = cp * cr;
    m[1][2] = -sp;
    m[2][0] = cy * sp * sr - sy * cr;
    m[2][1] = cy * sp * cr + sy * sr;
    m[2][2] = cy * cp;
}

void angles_rm(const double m[3][3], double& yaw, double& pitch, double& roll)
{
    if (m[1][2] > -1 && m[1][2] < 1) {
        pitch = std::asin(-m[1][2]);
        yaw = std::atan2(m[0][2], m[2][2]);
        roll = std::atan2(m[1][0], m[1][1]);
    } else {
        roll = 0;
        yaw = std::atan2(-m[2][0], m[0][0]);
        pitch = m[1][2] > 0 ? -kPi / 2 : kPi / 2;
    }
}

void rotate_one(Param& q, double yaw, double pitch, double roll)
{
    double r[3][3], c[3][3], m[3][3];
    rot_rm(r, yaw, pitch, roll);
    rot_rm(c, q.yaw * kRadPerDeg, q.pitch * kRadPerDeg, q.roll * kRadPerDeg);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            // cv::Matx product order: sum over k of r(i,k) c(k,j), k ascending
            double s = 0;
            for (int k = 0; k < 3; k++) s += r[i][k] * c[k][j];
            m[i][j] = s;
        }
    double y, p, rr;
    angles_rm(m, y, p, rr);
    q.yaw = y * kDegPerRad;
    q.pitch = p * kDegPerRad;
    q.roll = rr * kDegPerRad;
}

// --------------------------------------------------------------- export ----

std::string num(double v)
{
    std::ostringstream o;
    o << v;
    return o.str();
}

int export_xml(const char* path, const Param* ps, int n)
{
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f.is_open()) return PANO_ERR_IO;
    if (n == 0) {
        f << "<Root />\n";
        return f.good() ? PANO_OK : PANO_ERR_IO;
    }
    f << "<Root>\n";
    for (int i = 0; i < n; i++) {
        const Param& q = ps[i];
        const std::pair<const char*, std::string> rows[] = {
            {"IMAGETYPE", std::to_string(q.imageType)},
            {"CROPMODE", std::to_string(q.cropMode)},
            {"HFOV", num(q.hfov)},
            {"VFOV", num(q.vfov)},
            {"Alpha", num(q.alpha)},
            {"Beta", num(q.beta)},
            {"Gamma", num(q.gamma)},
            {"ShiftX", num(q.shiftX)},
            {"ShiftY", num(q.shiftY)},
            {"ShearX", num(q.shearX)},
            {"ShearY", num(q.shearY)},
            {"CROPLEFT", std::to_string(q.cropX)},
            {"CROPRIGHT", std::to_string(q.cropX + q.cropWidth)},
            {"CROPTOP", std::to_string(q.cropY)},
            {"CROPBOTTOM", std::to_string(q.cropY + q.cropHeight)},
            {"CircleX", std::to_string(q.circleX)},
            {"CircleY", std::to_string(q.circleY)},
            {"CircleR", std::to_string(q.circleR)},
            {"YAW", num(q.yaw)},
            {"PITCH", num(q.pitch)},
            {"ROLL", num(q.roll)},
        };
        f << "    <POSITION>\n";
        for (const auto& r : rows) f << "        <" << r.first << ">" << r.second << "</" << r.first << ">\n";
        f << "    </POSITION>\n";
    }
    f << "</Root>\n";
    return f.good() ? PANO_OK : PANO_ERR_IO;
}

bool args_ok(const char* path, const pano_photo_param* out, int cap)
{
    return path && cap >= 0 && (cap == 0 || out);
}

}  // namespace

extern "C" {

int pano_load_photo_params_xml(const char* path, pano_photo_param* out, int cap)
{
    if (!args_ok(path, out, cap)) return PANO_ERR_ARG;
    std::vector<Param> v;
    const int st = load_xml_params(path, v);
    return st < 0 ? st : hand_out(v, out, cap);
}

int pano_load_photo_params_pts(const char* path, pano_photo_param* out, int cap)
{
    if (!args_ok(path, out, cap)) return PANO_ERR_ARG;
    std::vector<Param> v;
    load_pts_params(path, v);
    return hand_out(v, out, cap);
}

int pano_load_photo_params(const char* path, pano_photo_param* out, int cap)
{
    if (!args_ok(path, out, cap)) return PANO_ERR_ARG;
    const char* dot = std::strrchr(path, '.');
    if (!dot) return PANO_ERR_ARG;   // "file does not have extention"
    return std::strcmp(dot + 1, "pts") == 0 ? pano_load_photo_params_pts(path, out, cap)
                                             : pano_load_photo_params_xml(path, out, cap);
}

int pano_export_photo_params_xml(const char* path, const pano_photo_param* params, int num)
{
    if (!path || num < 0 || (num > 0 && !params)) return PANO_ERR_ARG;
    return export_xml(path, params, num);
}

int pano_rotate_cameras(pano_photo_param* params, int num, double yaw, double pitch, double roll)
{
    if (num < 0 || (num > 0 && !params)) return PANO_ERR_ARG;
    for (int i = 0; i < num; i++) rotate_one(params[i], yaw, pitch, roll);
    return PANO_OK;
}

int pano_rotate_photo_params_xml(const char* src, const char* dst, double yaw, double pitch, double roll)
{
    if (!src || !dst) return PANO_ERR_ARG;
    std::vector<Param> v;
    const int st = load_xml_params(src, v);
    if (st < 0) return st;
    for (Param& q : v) rotate_one(q, yaw, pitch, roll);
    return export_xml(dst, v.data(), (int)v.size());
}

}  // extern "C"
