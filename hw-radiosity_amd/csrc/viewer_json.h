// The WebGL viewer's model file made from a .fin, byte for byte as the
// reference's converter writes it (WebGL_demo/model-converter/main.cpp:81-246,
// `tri2json -i <file.fin> -o <file.json>`).
#pragma once

#include <string>

namespace rgpu {

// Reads the .fin at fin_path and writes the viewer JSON to json_path: the
// vertices moved and scaled into a cube of side 10 centred on the origin
// (main.cpp:224-240), normals as read, colours / 255, three texture
// coordinates of 1 per vertex and indices 0 .. 3n-1.  Returns 0 and the
// triangle count and number of triangles whose scaled area is below 1e-6
// (the reference warns once for each), or -1 with err set for an unreadable
// input (the reference aborts) or an output that cannot be written.
int fin_to_json(const std::string &fin_path, const std::string &json_path, int &triangles, int &tiny, std::string &err);

}  // namespace rgpu
