// rt_svg.cpp -- SVG writers of the four models, the reference's drop-in
// drawing helpers (DominoTiler.cpp:640-816, LozengeTiler.cpp:558-693,
// RectTriangleTiler.cpp:646-776, TriangleDimerTiler.cpp:984-1117).
//
// The files are meant to be the reference's byte for byte: same element
// text, same order, numbers through the same iostream formatting (default
// precision, so e.g. 3.5 prints "3.5" and 1.5e6 prints "1.5e+06").  The
// element text is therefore the reference's format; the drawing logic is
// written here once per model around a small emitter.
#include <cmath>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "rt_internal.hpp"

namespace rt {
namespace {

constexpr double kSqrt3 = 1.7320508075688772935274;  // common.h:48

struct Svg {
  std::ostringstream o;

  // Reference headers: square viewBox, pixel size and optional
  // preserveAspectRatio="none".
  void open(const std::string& hw, double box, bool stretch) {
    o << "<svg xmlns='http://www.w3.org/2000/svg' xmlns:xlink=\"http://www.w3.org/1999/xlink\" "
      << hw << " viewBox=\"0 0 " << box << " " << box << "\"" << (stretch ? " preserveAspectRatio=\"none\"" : "")
      << ">\n";
  }
  void defs(const std::vector<std::string>& items) {
    o << "<defs>\n";
    for (const std::string& s : items) o << s;
    o << "</defs>\n";
  }
  // <use> in the spaced form most writers use ...
  template <class X, class Y>
  void use(const char* id, X x, Y y) {
    o << "<use xlink:href = \"#" << id << "\" x = \"" << x << "\" y = \"" << y << "\"/>\n";
  }
  // ... and the compact form of the domino dimer picture.
  void use_tight(const char* id, double x, double y) {
    o << "<use xlink:href=\"#" << id << "\" x=\"" << x << "\" y=\"" << y << "\"/>\n";
  }
  int save(const char* path) {
    o << "</svg>";
    std::ofstream f(path, std::ios::binary);
    if (!f) return RT_EIO;
    const std::string s = o.str();
    f.write(s.data(), (std::streamsize)s.size());
    return f ? RT_OK : RT_EIO;
  }
};

inline int isqrt(size_t n) { return (int)std::sqrt((double)n); }

// polygon / polyline group definitions
std::string group(const char* id, const std::string& body) {
  return std::string("<g id=\"") + id + "\">  " + body + " </g>\n";
}

const char* kHW500 = "height=\"500\" width=\"500\"";
const char* kHWHex = "height=\"433.013\" width=\"500\"";

}  // namespace

// ---------------------------------------------------------------- domino
int domino_domain_svg(const domain& d, const char* path) {
  const int M = isqrt(d.size());
  Svg s;
  s.open(kHW500, M, false);
  s.defs({group("b", "<polygon points = \"-.5,-.5 .5,-.5 .5,.5 -.5,.5\" fill=\"blue\"/>"),
          group("w", "<polygon points = \"-.5,-.5 .5,-.5 .5,.5 -.5,.5\" fill=\"grey\"/>")});
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < M; ++j)
      if (d[(size_t)i * M + j] == 1) s.use((i + j) % 2 == 0 ? "b" : "w", j + .5, i + .5);
  return s.save(path);
}

int domino_dimer_svg(const tiling& t, const domain& d, const char* path) {
  const int N = isqrt(t.size()), M = N - 1;
  auto D = [&](long k) { return k >= 0 && (size_t)k < d.size() ? d[(size_t)k] : 0; };
  Svg s;
  s.open(kHW500, M, false);
  s.defs({group("hd", "<polyline points= \"-.5,0 .5,0\" stroke=\"red\" stroke-width=\".05\"/>"),
          group("he", "<polyline points= \"-.5,0 .5,0\" stroke=\"black\" stroke-width=\".01\"/>"),
          group("vd", "<polyline points= \"0,-.5 0,.5\" stroke=\"red\" stroke-width=\".05\"/>"),
          group("ve", "<polyline points= \"0,-.5 0,.5\" stroke=\"black\" stroke-width=\".01\"/>"),
          "<g id=\"b\">  <circle cx=\"0\" cy=\"0\" r=\".1\" stroke=\"black\" stroke-width=\".02\" "
          "fill=\"black\"/> </g> \n",
          "<g id=\"w\">  <circle cx=\"0\" cy=\"0\" r=\".1\" stroke=\"black\" stroke-width=\".02\" "
          "fill=\"white\" /> </g>\n"});
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const int v = t[(size_t)i * N + j];
      if (v & 1) s.use_tight("hd", j, i - .5);
      if (v & 4) s.use_tight("vd", j - .5, i);
    }
  for (int i = 1; i < M; ++i)
    for (int j = 1; j < M; ++j) {
      const long k = (long)i * M + j;
      if (D(k) != 1) continue;
      const double x = j + .5, y = i + .5;
      if ((i + j) % 2 == 1) {
        if (D(k + 1) == 1) s.use_tight("he", x + .5, y);
        if (D(k - 1) == 1) s.use_tight("he", x - .5, y);
      }
      if (D(k - M) == 1) s.use_tight("ve", x, y - .5);
      if (D(k + M) == 1) s.use_tight("ve", x, y + .5);
    }
  for (int i = 1; i < M; ++i)
    for (int j = 1; j < M; ++j)
      if (D((long)i * M + j) == 1) s.use((i + j) % 2 == 0 ? "b" : "w", j + .5, i + .5);
  return s.save(path);
}

// TilingToSVG and MayaToSVG draw the same markers (horizontal bit 1,
// vertical bit 4, two colours by vertex parity) with different symbols.
static int domino_markers_svg(const tiling& t, const char* path, bool maya) {
  const int N = isqrt(t.size());
  Svg s;
  s.open(kHW500, N - 1, false);
  if (!maya) {
    const std::string h = "<polygon points = \"-1,0 -1,-1 1,-1 1,0\" fill=\"white\" stroke=\"black\" "
                          "stroke-width=\".05\"/>";
    const std::string v = "<polygon points = \"-1,-1 0,-1 0,1 -1,1\" fill=\"white\" stroke=\"black\" "
                          "stroke-width=\".05\"/>";
    s.o << "<defs>\n" << group("h1", h) << group("v1", v) << group("h2", h) << group("v2", v)
        << "</defs>\n";
  } else {
    auto disc = [](const char* cx, const char* cy, const char* fill) {
      return std::string("<circle cx=\"") + cx + "\" cy=\"" + cy +
             "\" r=\".4\" stroke=\"black\" stroke-width=\".02\" fill=\"" + fill + "\"/>";
    };
    auto pair = [&](const char* id, const char* cx2, const char* cy2, const char* fill) {
      return std::string("<g id=\"") + id + "\">" + disc("-.5", "-.5", fill) + " " +
             disc(cx2, cy2, fill) + "</g>\n";
    };
    s.o << "<defs>\n" << pair("h1", ".5", "-.5", "black") << pair("h2", ".5", "-.5", "white")
        << pair("v1", "-.5", ".5", "black") << pair("v2", "-.5", ".5", "white") << "</defs>\n";
  }
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const int v = t[(size_t)i * N + j];
      const bool odd = (i + j) % 2 == 1;
      if (v & 1) s.use(odd ? "h1" : "h2", (double)j, (double)i);
      if (v & 4) s.use(odd ? "v1" : "v2", (double)j, (double)i);
    }
  return s.save(path);
}

int domino_tiling_svg(const tiling& t, const char* path) { return domino_markers_svg(t, path, false); }
int domino_maya_svg(const tiling& t, const char* path) { return domino_markers_svg(t, path, true); }

// --------------------------------------------------------------- lozenge
int lozenge_domain_svg(const domain& d, const char* path) {
  const int M = isqrt(d.size() / 2);
  Svg s;
  s.open(kHWHex, 3 / 2.0 * M, true);
  s.o << "<defs>\n"
      << group("w", "<polygon points = \"0,1 .5,0 1,1\" fill=\"lightsteelblue\"/>")
      << group("b", "<polygon points = \"0,0 1,0 .5,1\" fill=\"slategrey\"/>") << "</defs>\n";
  for (int i = 0; i < M; i++)
    for (int j = 0; j < 2 * M; j++)
      if (d[(size_t)i * M * 2 + j] == 1) s.use(j % 2 == 0 ? "b" : "w", .5 * j + .5 * i, i);
  return s.save(path);
}

int lozenge_tiling_svg(const tiling& t, const char* path) {
  const int N = isqrt(t.size()), M = N - 1;
  Svg s;
  s.open(kHWHex, 3 / 2.0 * M, true);
  s.o << "<defs>\n"
      << group("r", "<polygon points = \"0,0 -1,0 -.5,-1 .5,-1\" fill=\"midnightblue\"/>")
      << group("g", "<polygon points = \"0,0 1,0 .5,-1 -.5,-1\" fill=\"lightsteelblue\"/>")
      << group("b", "<polygon points = \"0,0 .5,-1 1,0 .5,1\" fill=\"slategrey\"/>") << "</defs>\n";
  for (int i = 1; i < N - 1; i++)
    for (int j = 1; j < N - 1; j++) {
      const int v = t[(size_t)i * N + j];
      if (v & 1) s.use("r", j + .5 * i, i);
      if (v & 2) s.use("g", j + .5 * i, i);
      if (v & 8) s.use("b", j + .5 * i, i);
    }
  return s.save(path);
}

int lozenge_dimer_svg(const tiling& t, const domain& d, const char* path) {
  const int N = isqrt(t.size()), M = N - 1;
  auto D = [&](long k) { return k >= 0 && (size_t)k < d.size() ? d[(size_t)k] : 0; };
  auto edge = [](const char* id, const char* pts, const char* w) {
    return std::string("<g id=\"") + id + "\">  <polyline points= \"" + pts +
           "\" stroke=\"black\" stroke-width=\"" + w + "\"/> </g>\n";
  };
  const char* p1 = "-.5,-.33333333333 0,-.666666666666";
  const char* p2 = "0,-.666666666 .5,-.333333333333";
  const char* p8 = ".5,-.333333333333 .5,.333333333333";
  Svg s;
  s.open(kHWHex, 3 / 2.0 * M, true);
  s.o << "<defs>\n" << edge("e1", p1, ".001") << edge("e2", p2, ".001") << edge("e8", p8, ".001")
      << edge("d1", p1, ".025") << edge("d2", p2, ".025") << edge("d8", p8, ".025")
      << "<g id=\"b\">  <ellipse cx=\".5\" cy=\".333333333333\" rx=\".1\" ry=\".11547\" "
         "stroke=\"black\" stroke-width=\".02\" fill=\"black\"/> </g> \n"
      << "<g id=\"w\">  <ellipse cx=\".5\" cy=\".666666666666\" rx=\".1\"  ry=\".11547\" "
         "stroke=\"black\" stroke-width=\".02\" fill=\"white\" /> </g>\n"
      << "</defs>\n";
  for (int i = 1; i < N - 1; i++)
    for (int j = 1; j < N - 1; j++) {
      const int v = t[(size_t)i * N + j];
      if (v & 1) s.use("d1", j + .5 * i, i);
      if (v & 2) s.use("d2", j + .5 * i, i);
      if (v & 8) s.use("d8", j + .5 * i, i);
    }
  const long W = 2L * M;
  for (int i = 1; i < N - 1; i++)
    for (int j = 1; j < 2 * N - 1; j++) {
      const long k = i * W + j;
      if (D(k) != 1 || j % 2 != 0) continue;
      const double x = .5 + .5 * j + .5 * i;
      if (D(k - 1) == 1) s.use("e1", x, i + 1);
      if (D(k + 1) == 1) s.use("e2", x, i + 1);
      if (D(k + W - 1) == 1) s.use("e8", x, i + 1);
    }
  for (int i = 1; i < N - 1; i++)
    for (int j = 1; j < 2 * N - 1; j++)
      if (D(i * W + j) == 1) s.use(j % 2 == 0 ? "b" : "w", .5 * i + .5 * j, (double)i);
  return s.save(path);
}

// ------------------------------------------------------ rectangle-triangle
// One triangle face (TilePicture, RectTriangleTiler.cpp:684-776): corners
// P, Q, R; edge midpoints p, q, r opposite them.  The 4-bit face code says
// which sub-triangles are filled (colour 1..3) and which edges are drawn.
static void recttri_face(std::ostringstream& o, int code, int i, int j) {
  static const char* fills[4] = {
      "", "\" style=\"fill:midnightblue;stroke:midnightblue;stroke-width:.4\" />\n",
      "\" style=\"fill:slategrey;stroke:slategrey;stroke-width:.4\" />\n",
      "\" style=\"fill:lightsteelblue;stroke:lightsteelblue;stroke-width:.4\" />\n"};
  static const char* edge_style = " style=\"stroke:rgb(0,0,0);stroke-width:.5\" stroke-linecap = \"round\" />\n";
  // Per code: whole-face colour, the six half-face colours
  // (pPR, pPQ, rRQ, rRP, qQR, qQP), full edges PQ, PR, QR, half edges p, q, r.
  struct Face {
    int whole, pPR, pPQ, rRQ, rRP, qQR, qQP;
    bool PQ, PR, QR, p, q, r;
  };
  static const Face faces[16] = {
      {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, {1, 0, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 1},
      {3, 0, 0, 0, 0, 0, 0, 1, 0, 1, 0, 1, 0}, {2, 0, 0, 0, 0, 0, 0, 1, 1, 0, 1, 0, 0},
      {2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1}, {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0},
      {3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0}, {2, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0},
      {1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0}, {3, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0},
      {0, 0, 0, 2, 1, 0, 0, 0, 1, 0, 0, 0, 1}, {0, 0, 0, 0, 0, 3, 1, 0, 0, 1, 0, 1, 0},
      {0, 3, 2, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0}, {0, 0, 0, 1, 2, 0, 0, 0, 0, 1, 0, 0, 1},
      {0, 0, 0, 0, 0, 1, 3, 1, 0, 0, 0, 1, 0}, {0, 2, 3, 0, 0, 0, 0, 0, 1, 0, 1, 0, 0}};
  const Face& f = faces[code & 15];
  auto vx = [](int a, int b) { return 20 * (a * 1 / 2.0 + b); };
  auto vy = [](int a) { return 20 * (a * kSqrt3 / 2.0); };
  double Px, Py, Qx, Qy, Rx, Ry;
  if (j % 2 == 1) {
    Px = vx(i + 1, j / 2), Py = vy(i + 1);
    Qx = vx(i + 1, j / 2 + 1), Qy = vy(i + 1);
    Rx = vx(i, j / 2 + 1), Ry = vy(i);
  } else {
    Px = vx(i, j / 2 + 1), Py = vy(i);
    Qx = vx(i, j / 2), Qy = vy(i);
    Rx = vx(i + 1, j / 2), Ry = vy(i + 1);
  }
  const double px = (Rx + Qx) / 2, py = (Ry + Qy) / 2;
  const double qx = (Px + Rx) / 2, qy = (Py + Ry) / 2;
  const double rx = (Px + Qx) / 2, ry = (Py + Qy) / 2;
  auto tri = [&](int colour, double ax, double ay, double bx, double by, double cx, double cy) {
    if (colour > 0)
      o << "<polygon points=\"" << ax << "," << ay << " " << bx << "," << by << " " << cx << ","
        << cy << " " << fills[colour];
  };
  tri(f.whole, Px, Py, Qx, Qy, Rx, Ry);
  tri(f.pPR, Px, Py, px, py, Rx, Ry);
  tri(f.pPQ, Px, Py, px, py, Qx, Qy);
  tri(f.rRQ, Rx, Ry, rx, ry, Qx, Qy);
  tri(f.rRP, Rx, Ry, rx, ry, Px, Py);
  tri(f.qQR, Qx, Qy, qx, qy, Rx, Ry);
  tri(f.qQP, Qx, Qy, qx, qy, Px, Py);
  auto line = [&](bool on, double ax, double ay, double bx, double by) {
    if (on)
      o << "<line x1=\"" << ax << "\" y1=\"" << ay << "\" x2=\"" << bx << "\" y2=\"" << by << "\""
        << edge_style;
  };
  line(f.PQ, Px, Py, Qx, Qy);
  line(f.PR, Px, Py, Rx, Ry);
  line(f.QR, Rx, Ry, Qx, Qy);
  line(f.p, Px, Py, px, py);
  line(f.q, Qx, Qy, qx, qy);
  line(f.r, Rx, Ry, rx, ry);
}

int recttri_tiling_svg(const tiling& t, const char* path) {
  const int N = isqrt(t.size()), M = N - 1;
  const double W = kSqrt3 * 2 * (N + 1), H = 3 / 2.0 * (N + 1), sc = 20;
  std::ostringstream o;
  o << "<svg xmlns='http://www.w3.org/2000/svg' height=\"" << sc * H << "\" width=\"" << sc * W
    << "\" >\n";
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      recttri_face(o, (t[(size_t)i * N + j + 1] & 0xf0) >> 4, i, 2 * j + 1);
      recttri_face(o, t[(size_t)i * N + j] & 0xf, i, 2 * j);
    }
  o << "</svg>";
  std::ofstream f(path, std::ios::binary);
  if (!f) return RT_EIO;
  const std::string s = o.str();
  f.write(s.data(), (std::streamsize)s.size());
  return f ? RT_OK : RT_EIO;
}

// ---------------------------------------------------------------- bibone
int bibone_domain_svg(const domain& d, const char* path) {
  const int M = isqrt(d.size() / 2);
  Svg s;
  s.open(kHWHex, 3 / 2.0 * M, true);
  s.o << "<defs>\n"
      << group("w", "<polygon points = \"0,1 .5,0 1,1\" fill=\"paleturquoise\"/>")
      << group("b", "<polygon points = \"0,0 1,0 .5,1\" fill=\"deepskyblue\"/>") << "</defs>\n";
  for (int i = 0; i < M; i++)
    for (int j = 0; j < 2 * M; j++)
      if (d[(size_t)i * M * 2 + j] == 1) s.use(j % 2 == 0 ? "b" : "w", .5 * j + .5 * i, i);
  return s.save(path);
}

int bibone_tiling_svg(const tiling& t, const char* path) {
  const int N = isqrt(t.size()), M = N / 2;
  auto shape = [](const char* id, const char* pts, const char* fill) {
    return std::string("<g id=\"") + id + "\">  <polygon points = \"" + pts + "\" fill=\"" + fill +
           "\" stroke=\"black\" stroke-width=\".05\"/></g>\n";
  };
  Svg s;
  s.open(kHWHex, 3 / 2.0 * M, true);
  s.o << "<defs>\n"
      << shape("dR",
               "1,-0.6666666667 0.5,-0.3333333333 0.5,0.3333333333 0,0.6666666667 0,1.333333333 "
               "0.5,1.666666667 1,1.333333333 1,0.6666666667 1.5,0.3333333333 1.5,-0.3333333333",
               "midnightblue")
      << shape("dL",
               "0,-0.6666666667 0.5,-0.3333333333 0.5,0.3333333333 1,0.6666666667 1,1.333333333 "
               "0.5,1.666666667 0,1.333333333 0,0.6666666667 -0.5,0.3333333333 -0.5,-0.3333333333",
               "lightsteelblue")
      << shape("dH",
               "0,-0.6666666667 0.5,-0.3333333333 1,-0.6666666667 1.5,-0.3333333333 "
               "1.5,0.3333333333 1,0.6666666667 0.5,0.3333333333 0,0.6666666667 -0.5,0.3333333333 "
               "-0.5,-0.3333333333",
               "slategrey")
      << "</defs>\n";
  for (int i = 0; i < M - 1; ++i)
    for (int j = 1; j < 2 * M - 1; j += 2) {
      const double x = .5 * j + .5 * i;
      const int h = t[(size_t)2 * i * N + j], r = t[(size_t)(2 * i + 1) * N + j];
      if (h & 4) s.use("dR", x, i);
      if (h & 8) s.use("dL", x, i);
      if (r & 1) s.use("dH", x, i);
    }
  return s.save(path);
}

int bibone_dimer_svg(const tiling& t, const domain& d, const char* path) {
  const int N = isqrt(t.size()), M = isqrt(d.size() / 2);
  auto D = [&](long k) { return k >= 0 && (size_t)k < d.size() ? d[(size_t)k] : 0; };
  auto stroke = [](const char* id, const char* pts, const char* colour, const char* w) {
    return std::string("<g id=\"") + id + "\">  <polyline points= \"" + pts + "\" stroke=\"" +
           colour + "\" stroke-width=\"" + w + "\"/> </g>\n";
  };
  auto dot = [](const char* id, const char* cx, const char* cy) {
    return std::string("<g id=\"") + id + "\">  <ellipse cx=\"" + cx + "\" cy=\"" + cy +
           "\" rx=\".1\" ry=\".1\" stroke=\"black\" stroke-width=\".02\" fill=\"black\"/> </g> \n";
  };
  Svg s;
  s.open(kHWHex, 3 / 2.0 * M, true);
  s.o << "<defs>\n"
      << stroke("dR", ".5,1 1,0", "darkblue", ".05") << stroke("dL", ".5,1 0,0", "darkred", ".05")
      << stroke("dH", "1,0 0,0", "darkgreen", ".05") << stroke("eR", ".5,1 1,0", "black", ".001")
      << stroke("eL", ".5,1 0,0", "black", ".001") << stroke("eH", "1,0 0,0", "black", ".001")
      << dot("v1", ".5", "0") << dot("v2", "1.5", "0") << dot("v3", "1", "1") << "</defs>\n";
  const long W = 2L * M;
  for (int i = 0; i < M - 1; ++i)
    for (int j = 1; j < 2 * M - 1; j += 2) {
      const double x = .5 * j + .5 * i;
      const int h = t[(size_t)2 * i * N + j], r = t[(size_t)(2 * i + 1) * N + j];
      const long k = i * W + j;
      if (h & 4) s.use("dR", x, i);
      else if (D(k - 1) == 1) s.use("eR", x, i);
      if (h & 8) s.use("dL", x, i);
      else if (D(k - 1) == 1) s.use("eL", x, i);
      if (r & 1) s.use("dH", x, i);
      else if (D(k - 1) == 1 || D(k - W) == 1) s.use("eH", x, i);
      if ((h & 4) == 0 && D(k) == 1) s.use("eR", x, i);
      if ((h & 8) == 0 && j > 1 && D(k - 2) == 1) s.use("eL", x, i);
    }
  for (int i = 0; i < M; i++)
    for (int j = 2; j < 2 * M; j += 2) {
      const long k = i * W + j;
      const double x = .5 * j + .5 * i;
      if (D(k) == 1 || D(k - 1) == 1) s.use("v1", x, i);
      if (D(k) == 1 || D(k + 1) == 1) s.use("v2", x, i);
      if (D(k) == 1 || D(k + 1) == 1 || D(k - 1) == 1) s.use("v3", x, i);
    }
  return s.save(path);
}

}  // namespace rt

// ------------------------------------------------------------------ C ABI
using namespace rt;

static std::vector<int> vec_of(const int32_t* p, size_t n) { return std::vector<int>(p, p + n); }

#define RT_SVG_CHECK(p, n)                        \
  if (!(p) || !path || (n) == 0) return RT_EINVAL;

extern "C" {

int rt_domino_domain_svg(const int32_t* d, size_t nd, const char* path) {
  RT_SVG_CHECK(d, nd);
  return domino_domain_svg(vec_of(d, nd), path);
}
int rt_domino_tiling_svg(const int32_t* t, size_t nt, const char* path) {
  RT_SVG_CHECK(t, nt);
  return domino_tiling_svg(vec_of(t, nt), path);
}
int rt_domino_maya_svg(const int32_t* t, size_t nt, const char* path) {
  RT_SVG_CHECK(t, nt);
  return domino_maya_svg(vec_of(t, nt), path);
}
int rt_domino_dimer_svg(const int32_t* t, size_t nt, const int32_t* d, size_t nd,
                        const char* path) {
  RT_SVG_CHECK(t, nt);
  const tiling tt = vec_of(t, nt);
  return domino_dimer_svg(tt, d ? vec_of(d, nd) : domino_tiling_to_domain(tt), path);
}
int rt_lozenge_domain_svg(const int32_t* d, size_t nd, const char* path) {
  RT_SVG_CHECK(d, nd);
  return lozenge_domain_svg(vec_of(d, nd), path);
}
int rt_lozenge_tiling_svg(const int32_t* t, size_t nt, const char* path) {
  RT_SVG_CHECK(t, nt);
  return lozenge_tiling_svg(vec_of(t, nt), path);
}
int rt_lozenge_dimer_svg(const int32_t* t, size_t nt, const int32_t* d, size_t nd,
                         const char* path) {
  RT_SVG_CHECK(t, nt);
  const tiling tt = vec_of(t, nt);
  return lozenge_dimer_svg(tt, d ? vec_of(d, nd) : lozenge_tiling_to_domain(tt), path);
}
int rt_recttri_tiling_svg(const int32_t* t, size_t nt, const char* path) {
  RT_SVG_CHECK(t, nt);
  return recttri_tiling_svg(vec_of(t, nt), path);
}
int rt_bibone_domain_svg(const int32_t* d, size_t nd, const char* path) {
  RT_SVG_CHECK(d, nd);
  return bibone_domain_svg(vec_of(d, nd), path);
}
int rt_bibone_tiling_svg(const int32_t* t, size_t nt, const char* path) {
  RT_SVG_CHECK(t, nt);
  return bibone_tiling_svg(vec_of(t, nt), path);
}
int rt_bibone_dimer_svg(const int32_t* t, size_t nt, const int32_t* d, size_t nd,
                        const char* path) {
  RT_SVG_CHECK(t, nt);
  const tiling tt = vec_of(t, nt);
  const int N = isqrt(nt);
  return bibone_dimer_svg(tt, d ? vec_of(d, nd) : bibone_tiling_to_domain(tt, N), path);
}

}  // extern "C"
