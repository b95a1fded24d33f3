// Graph update streams (the GraphStream / BinaryFileStream interface the reference's tools and
// driver use, include/graph_sketch_driver.h:122-176, tools/cuda_process_stream.cu:82-84; the
// implementation lives in the un-vendored StreamingUtilities dependency and is restated here).
// Binary format: 12-byte header {u32 num_vertices, u64 num_updates}, then 9-byte records
// {u8 type (0 insert, 1 delete), u32 src, u32 dst}, little-endian, packed.
#pragma once
#include <atomic>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <string>

#include "types.h"

static constexpr edge_id_t END_OF_STREAM = (edge_id_t)-1;

class GraphStream {
 public:
  virtual ~GraphStream() {}
  node_id_t vertices() const { return num_vertices; }
  edge_id_t edges() const { return num_edges; }
  // Fills up to num_updates updates; returns how many. At the break point (or the end of the
  // stream) it returns a single update of type BREAKPOINT.
  virtual size_t get_update_buffer(GraphStreamUpdate* upd_buf, size_t num_updates) = 0;
  // Updates [current position, break_idx) are returned before the next BREAKPOINT; false if the
  // stream is already past break_idx.
  virtual bool set_break_point(edge_id_t break_idx) = 0;
  virtual void serialize_metadata(std::ostream& out) = 0;
  bool get_update_is_thread_safe() const { return update_is_thread_safe; }

 protected:
  node_id_t num_vertices = 0;
  edge_id_t num_edges = 0;
  bool update_is_thread_safe = false;
};

class BinaryFileStream : public GraphStream {
 public:
  // open_read = false creates (truncates) the file for writing
  explicit BinaryFileStream(std::string file_name, bool open_read = true);
  ~BinaryFileStream() override;
  BinaryFileStream(const BinaryFileStream&) = delete;
  BinaryFileStream& operator=(const BinaryFileStream&) = delete;

  size_t get_update_buffer(GraphStreamUpdate* upd_buf, size_t num_updates) override;
  bool set_break_point(edge_id_t break_idx) override;
  void serialize_metadata(std::ostream& out) override;

  void write_header(node_id_t num_verts, edge_id_t num_edg);
  void write_updates(const GraphStreamUpdate* upd, edge_id_t num_updates);

  static constexpr size_t header_size = sizeof(node_id_t) + sizeof(edge_id_t);
  static constexpr size_t edge_size = sizeof(uint8_t) + 2 * sizeof(node_id_t);

 private:
  int fd = -1;
  bool read_mode;
  std::string file_name;
  std::atomic<edge_id_t> stream_off{0};  // next record to hand out
  std::atomic<edge_id_t> break_idx{0};
  edge_id_t write_off = 0;
};

// Text stream (the reference's tests read their res/*.txt graphs with StreamingUtilities'
// AsciiFileStream, test/cc_alg_test.cpp:46): a "num_vertices num_edges" line, then one update per
// line, "src dst" or, with has_type, "type src dst" (0 insert, 1 delete). Read by one thread.
class AsciiFileStream : public GraphStream {
 public:
  explicit AsciiFileStream(std::string file_name, bool has_type = true);
  AsciiFileStream(const AsciiFileStream&) = delete;
  AsciiFileStream& operator=(const AsciiFileStream&) = delete;

  size_t get_update_buffer(GraphStreamUpdate* upd_buf, size_t num_updates) override;
  bool set_break_point(edge_id_t break_idx) override;
  void serialize_metadata(std::ostream& out) override;

 private:
  std::string file_name;
  bool has_type;
  std::ifstream in;
  edge_id_t stream_off = 0;
  edge_id_t break_idx = 0;
};
