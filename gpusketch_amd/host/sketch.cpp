// Host Sketch (src/sketch.cpp of the reference: constructors :8-64, update :66-99,
// zero_contents :102-108, sample :110-132, exhaustive_sample :134-163, merges :165-202,
// serialize :204-206, operators :208-243).
#include "gpusketch/sketch.h"

#include <cstring>

void Sketch::init_geometry(vec_t vector_len, size_t samples, size_t cols) {
  num_samples = samples;
  cols_per_sample = cols;
  num_columns = num_samples * cols_per_sample;
  bkt_per_col = calc_bkt_per_col(vector_len);
  num_buckets = num_columns * bkt_per_col + 1;  // + the deterministic bucket
}

Sketch::Sketch(vec_t vector_len, uint64_t seed, node_id_t sketch_id, Bucket* _buckets,
               size_t samples, size_t cols)
    : seed(seed) {
  init_geometry(vector_len, samples, cols);
  buckets = _buckets + (size_t)sketch_id * num_buckets;
  borrowed = true;
  std::memset(buckets, 0, bucket_array_bytes());
}

Sketch::Sketch(vec_t vector_len, uint64_t seed, size_t samples, size_t cols) : seed(seed) {
  init_geometry(vector_len, samples, cols);
  buckets = new Bucket[num_buckets]();
}

Sketch::Sketch(vec_t vector_len, uint64_t seed, std::istream& binary_in, size_t samples,
               size_t cols)
    : seed(seed) {
  init_geometry(vector_len, samples, cols);
  buckets = new Bucket[num_buckets]();
  binary_in.read(reinterpret_cast<char*>(buckets), bucket_array_bytes());
}

Sketch::Sketch(const Sketch& s)
    : seed(s.seed),
      num_samples(s.num_samples),
      cols_per_sample(s.cols_per_sample),
      num_columns(s.num_columns),
      bkt_per_col(s.bkt_per_col),
      num_buckets(s.num_buckets) {
  buckets = new Bucket[num_buckets];
  std::memcpy(buckets, s.buckets, bucket_array_bytes());
}

Sketch::~Sketch() {
  if (!borrowed) delete[] buckets;
}

void Sketch::update(const vec_t update_idx) {
  const vec_hash_t checksum = Bucket_Boruvka::get_index_hash(update_idx, checksum_seed());
  Bucket_Boruvka::update(buckets[num_buckets - 1], update_idx, checksum);
  for (size_t c = 0; c < num_columns; ++c) {
    const col_hash_t depth =
        Bucket_Boruvka::get_index_depth(update_idx, column_seed(c), bkt_per_col);
    if (depth >= bkt_per_col) continue;
#ifdef L0_SAMPLING
    for (col_hash_t d = 0; d <= depth; ++d)  // CubeSketch: every depth up to the sampled one
      Bucket_Boruvka::update(buckets[c * bkt_per_col + d], update_idx, checksum);
#else
    Bucket_Boruvka::update(buckets[c * bkt_per_col + depth], update_idx, checksum);
#endif
  }
}

void Sketch::zero_contents() {
  std::memset(buckets, 0, bucket_array_bytes());
  reset_sample_state();
}

SketchSample Sketch::sample() {
  if (sample_idx >= num_samples) throw OutOfSamplesException(seed, num_samples, sample_idx);
  const size_t first_column = (sample_idx++) * cols_per_sample;
  const Bucket& det = buckets[num_buckets - 1];
  if (det.alpha == 0 && det.gamma == 0) return {0, ZERO};
  if (Bucket_Boruvka::is_good(det, checksum_seed())) return {det.alpha, GOOD};
  for (size_t c = first_column; c < first_column + cols_per_sample; ++c)
    for (size_t d = 0; d < bkt_per_col; ++d) {
      const Bucket& b = buckets[c * bkt_per_col + d];
      if (Bucket_Boruvka::is_good(b, checksum_seed())) return {b.alpha, GOOD};
    }
  return {0, FAIL};
}

ExhaustiveSketchSample Sketch::exhaustive_sample() {
  if (sample_idx >= num_samples) throw OutOfSamplesException(seed, num_samples, sample_idx);
  const size_t first_column = (sample_idx++) * cols_per_sample;
  std::unordered_set<vec_t> found;
  const Bucket& det = buckets[num_buckets - 1];
  if (det.alpha == 0 && det.gamma == 0) return {found, ZERO};
  if (Bucket_Boruvka::is_good(det, checksum_seed())) {
    found.insert(det.alpha);
    return {found, GOOD};
  }
  for (size_t c = first_column; c < first_column + cols_per_sample; ++c)
    for (size_t d = 0; d < bkt_per_col; ++d) {
      const Bucket& b = buckets[c * bkt_per_col + d];
      if (Bucket_Boruvka::is_good(b, checksum_seed())) found.insert(b.alpha);
    }
  return {found, found.empty() ? FAIL : GOOD};
}

void Sketch::merge(const Sketch& other) { merge_raw_bucket_buffer(other.buckets); }

void Sketch::range_merge(const Sketch& other, size_t start_sample, size_t n_samples) {
  if (start_sample + n_samples > num_samples) {
    sample_idx = num_samples;  // out of range: the sketch can no longer be sampled
    return;
  }
  sample_idx = std::max(sample_idx, start_sample);
  Bucket_Boruvka::update(buckets[num_buckets - 1], other.buckets[num_buckets - 1].alpha,
                         other.buckets[num_buckets - 1].gamma);
  const size_t b0 = start_sample * cols_per_sample * bkt_per_col;
  const size_t nb = n_samples * cols_per_sample * bkt_per_col;
  for (size_t i = b0; i < b0 + nb; i++)
    Bucket_Boruvka::update(buckets[i], other.buckets[i].alpha, other.buckets[i].gamma);
}

void Sketch::merge_raw_bucket_buffer(const Bucket* raw) {
  for (size_t i = 0; i < num_buckets; i++)
    Bucket_Boruvka::update(buckets[i], raw[i].alpha, raw[i].gamma);
}

void Sketch::serialize(std::ostream& binary_out) const {
  binary_out.write(reinterpret_cast<const char*>(buckets), bucket_array_bytes());
}

bool operator==(const Sketch& a, const Sketch& b) {
  return a.num_buckets == b.num_buckets && a.seed == b.seed &&
         std::memcmp(a.buckets, b.buckets, a.bucket_array_bytes()) == 0;
}

std::ostream& operator<<(std::ostream& os, const Sketch& s) {
  auto show = [&](const Bucket& b) {
    os << " a:" << b.alpha << " c:" << b.gamma
       << (Bucket_Boruvka::is_good(b, s.checksum_seed()) ? " good" : " bad") << std::endl;
  };
  show(s.buckets[s.num_buckets - 1]);
  for (size_t c = 0; c < s.num_columns; ++c) {
    for (size_t d = 0; d < s.bkt_per_col; ++d) show(s.buckets[c * s.bkt_per_col + d]);
    os << std::endl;
  }
  return os;
}
