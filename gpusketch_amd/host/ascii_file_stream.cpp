// AsciiFileStream reader (format in graph_stream.h).
#include <fstream>
#include <stdexcept>

#include "gpusketch/graph_stream.h"

AsciiFileStream::AsciiFileStream(std::string name, bool typed)
    : file_name(std::move(name)), has_type(typed), in(file_name) {
  if (!in) throw std::runtime_error("AsciiFileStream: cannot open " + file_name);
  uint64_t n = 0, m = 0;
  in >> n >> m;
  if (!in) throw std::runtime_error("AsciiFileStream: bad header in " + file_name);
  num_vertices = (node_id_t)n;
  num_edges = m;
  break_idx = m;
  update_is_thread_safe = false;
}

size_t AsciiFileStream::get_update_buffer(GraphStreamUpdate* buf, size_t want) {
  if (want == 0) return 0;
  std::ifstream& f = in;
  const edge_id_t limit = std::min<edge_id_t>(break_idx, num_edges);
  if (stream_off >= limit) {
    buf[0].type = BREAKPOINT;
    buf[0].edge = {0, 0};
    return 1;
  }
  size_t got = 0;
  while (got < want && stream_off < limit) {
    uint64_t t = 0, a = 0, b = 0;
    if (has_type) f >> t;
    f >> a >> b;
    if (!f) throw std::runtime_error("AsciiFileStream: truncated stream " + file_name);
    buf[got].type = t ? DELETE : INSERT;
    buf[got].edge = {(node_id_t)a, (node_id_t)b};
    got++;
    stream_off++;
  }
  return got;
}

bool AsciiFileStream::set_break_point(edge_id_t idx) {
  if (idx < stream_off) return false;
  break_idx = idx;
  return true;
}

void AsciiFileStream::serialize_metadata(std::ostream& out) {
  out << "AsciiFileStream " << file_name << " " << stream_off << std::endl;
}
