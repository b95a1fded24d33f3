// BinaryFileStream reader and writer (format in graph_stream.h). Readers share one atomic cursor
// and read with pread, so several stream threads may call get_update_buffer at once.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "gpusketch/graph_stream.h"

BinaryFileStream::BinaryFileStream(std::string name, bool open_read)
    : read_mode(open_read), file_name(std::move(name)) {
  fd = open_read ? open(file_name.c_str(), O_RDONLY)
                 : open(file_name.c_str(), O_RDWR | O_CREAT | O_TRUNC, 0644);
  if (fd < 0) throw std::runtime_error("BinaryFileStream: cannot open " + file_name);
  if (open_read) {
    unsigned char h[header_size];
    if (pread(fd, h, header_size, 0) != (ssize_t)header_size)
      throw std::runtime_error("BinaryFileStream: short header in " + file_name);
    std::memcpy(&num_vertices, h, sizeof(node_id_t));
    std::memcpy(&num_edges, h + sizeof(node_id_t), sizeof(edge_id_t));
    break_idx = num_edges;
  }
  update_is_thread_safe = true;
}

BinaryFileStream::~BinaryFileStream() {
  if (fd >= 0) close(fd);
}

size_t BinaryFileStream::get_update_buffer(GraphStreamUpdate* buf, size_t want) {
  if (!read_mode) throw std::runtime_error("BinaryFileStream: not opened for reading");
  if (want == 0) return 0;
  const edge_id_t limit = std::min<edge_id_t>(break_idx.load(), num_edges);
  edge_id_t cur = stream_off.load();
  edge_id_t take;
  do {
    if (cur >= limit) {
      buf[0].type = BREAKPOINT;
      buf[0].edge = {0, 0};
      return 1;
    }
    take = std::min<edge_id_t>(want, limit - cur);
  } while (!stream_off.compare_exchange_weak(cur, cur + take));
  std::vector<unsigned char> raw(take * edge_size);
  const off_t off = (off_t)(header_size + cur * edge_size);
  size_t got = 0;
  while (got < raw.size()) {
    ssize_t r = pread(fd, raw.data() + got, raw.size() - got, off + (off_t)got);
    if (r <= 0) throw std::runtime_error("BinaryFileStream: short read in " + file_name);
    got += (size_t)r;
  }
  for (edge_id_t i = 0; i < take; i++) {
    const unsigned char* p = raw.data() + i * edge_size;
    buf[i].type = p[0] ? DELETE : INSERT;
    std::memcpy(&buf[i].edge.src, p + 1, sizeof(node_id_t));
    std::memcpy(&buf[i].edge.dst, p + 1 + sizeof(node_id_t), sizeof(node_id_t));
  }
  return take;
}

bool BinaryFileStream::set_break_point(edge_id_t idx) {
  if (idx < stream_off.load()) return false;
  break_idx = idx;
  return true;
}

void BinaryFileStream::serialize_metadata(std::ostream& out) {
  out << "BinaryFileStream " << file_name << " " << stream_off.load() << std::endl;
}

void BinaryFileStream::write_header(node_id_t nv, edge_id_t ne) {
  unsigned char h[header_size];
  std::memcpy(h, &nv, sizeof(node_id_t));
  std::memcpy(h + sizeof(node_id_t), &ne, sizeof(edge_id_t));
  if (pwrite(fd, h, header_size, 0) != (ssize_t)header_size)
    throw std::runtime_error("BinaryFileStream: header write failed");
  num_vertices = nv;
  num_edges = ne;
}

void BinaryFileStream::write_updates(const GraphStreamUpdate* upd, edge_id_t count) {
  std::vector<unsigned char> raw(count * edge_size);
  for (edge_id_t i = 0; i < count; i++) {
    unsigned char* p = raw.data() + i * edge_size;
    p[0] = upd[i].type == DELETE ? 1 : 0;
    std::memcpy(p + 1, &upd[i].edge.src, sizeof(node_id_t));
    std::memcpy(p + 1 + sizeof(node_id_t), &upd[i].edge.dst, sizeof(node_id_t));
  }
  const off_t off = (off_t)(header_size + write_off * edge_size);
  if (pwrite(fd, raw.data(), raw.size(), off) != (ssize_t)raw.size())
    throw std::runtime_error("BinaryFileStream: write failed");
  write_off += count;
}
