// One rotating update (mh_scene_update's device pass) run on the host, one lane at a time,
// from the kernels' own device functions: the leaf pass, the leaf splits, the rotations
// depth by depth (deepest first) and the PairNodes of the new tree. Built with
// AddressSanitizer over buffers of exact size by tests/test_update_host_cpu.py.
//
// usage: update_host DIR NEXT_FREE
//   DIR holds nodes.bin (RefNode[2T]), ids.bin, verts.bin, idx.bin, costs.bin, ssize.bin;
//   writes nodes_out.bin, ids_out.bin and prints "records R leaves L depth D".
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

dim3 threadIdx, blockIdx, gridDim, blockDim;
#include "rotate.hip"

using namespace mh;

template <class T>
static std::vector<T> load(const std::string& f) {
  FILE* fp = std::fopen(f.c_str(), "rb");
  if (!fp) std::exit(2);
  std::fseek(fp, 0, SEEK_END);
  const long n = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  std::vector<T> v(n / sizeof(T));
  if (std::fread(v.data(), 1, n, fp) != (size_t)n) std::exit(2);
  std::fclose(fp);
  return v;
}
// A heap copy of exactly v.size() elements, so that any access past the end is reported.
template <class T>
static T* exact(const std::vector<T>& v, size_t n = 0) {
  T* p = new T[n ? n : v.size()]();
  std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}
template <class T>
static void save(const std::string& f, const T* p, size_t n) {
  FILE* fp = std::fopen(f.c_str(), "wb");
  std::fwrite(p, sizeof(T), n, fp);
  std::fclose(fp);
}

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  const std::string d = argv[1];
  const auto nodes = load<RefNode>(d + "/nodes.bin");
  const auto ids = load<int32_t>(d + "/ids.bin");
  const int T = (int)ids.size(), nrec = (int)nodes.size();
  DevUpdate u{};
  u.refs = exact(nodes);
  u.ids = exact(ids);
  u.verts = exact(load<float>(d + "/verts.bin"));
  u.idx = exact(load<uint32_t>(d + "/idx.bin"));
  u.costs = exact(load<float>(d + "/costs.bin"));
  u.ssize = exact(load<uint32_t>(d + "/ssize.bin"));
  u.pairs = new PairNode[nrec / 2]();
  u.big_count = new int32_t[T]();
  u.parent = new int32_t[nrec]();
  u.split_list = new int32_t[T]();
  u.leaf_at = new int32_t[T];
  unsigned* ctr = new unsigned[3];
  ctr[0] = (unsigned)std::atoi(argv[2]);
  ctr[1] = ctr[2] = 0;
  u.next_free = ctr;
  u.err = reinterpret_cast<int*>(ctr + 1);
  u.split_count = ctr + 2;
  u.n_records = nrec;
  u.n_objects = T;
  u.large = (uint32_t)(nrec / 20);
  // the level lists and leaves of the current tree, and the leaf starting at each slot
  std::vector<std::vector<int>> levels;
  std::vector<int> leaves;
  std::vector<std::pair<int, int>> st{{0, 0}};
  while (!st.empty()) {
    const auto [r, dep] = st.back();
    st.pop_back();
    if (u.refs[r].children) {
      leaves.push_back(r);
      continue;
    }
    if ((int)levels.size() <= dep) levels.resize(dep + 1);
    levels[dep].push_back(r);
    st.push_back({u.refs[r].child + 1, dep + 1});
    st.push_back({u.refs[r].child, dep + 1});
  }
  for (int i = 0; i < T; ++i) u.leaf_at[i] = -1;
  for (int r : leaves) u.leaf_at[u.refs[r].child] = r;
  if (levels.empty()) return 3;  // the update's device pass needs an inner root
  for (int i = 0; i < T; ++i) {  // k_upd_leaves, lane i
    blockIdx.x = i / 256;
    threadIdx.x = i % 256;
    k_upd_leaves(u, T);
  }
  gridDim.x = 1;
  blockIdx.x = 0;
  for (unsigned t = 0; t < 256; ++t) {  // k_split_leaves, every lane of one workgroup
    threadIdx.x = t;
    k_split_leaves(u);
  }
  if (ctr[1]) return 4;
  for (int dep = (int)levels.size() - 1; dep >= 0; --dep)  // k_upd_level / k_upd_top
    for (int r : levels[dep]) upd_node(u, r);
  // the new tree: PairNodes and leaf starts, as the level pass stores them
  std::vector<std::pair<int, int>> q{{0, 0}};
  int depth = 0;
  size_t n_leaves = 0, n_records = 0;
  while (!q.empty()) {
    const auto [r, dep] = q.back();
    q.pop_back();
    ++n_records;
    depth = dep > depth ? dep : depth;
    const RecWords w = load_rec(u.refs, r);
    if (rec_children(w) != 0) {
      ++n_leaves;
      u.leaf_at[rec_child(w)] = r;
      continue;
    }
    const int c = rec_child(w);
    store_pair(u, r, c, rec_word(w), load_rec(u.refs, c), load_rec(u.refs, c + 1));
    q.push_back({c, dep + 1});
    q.push_back({c + 1, dep + 1});
  }
  save(d + "/nodes_out.bin", u.refs, (size_t)nrec);
  save(d + "/ids_out.bin", u.ids, (size_t)T);
  std::printf("records %zu leaves %zu depth %d\n", n_records, n_leaves, depth);
  return 0;
}
