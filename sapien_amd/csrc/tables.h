// Device-resident model tables (one contiguous block per model and precision).
//
// Layout in HBM: a 256-byte header followed by nj JointRec and nlinks LinkRec,
// all in the internal depth-first joint order (parents before children). Every
// kernel reads these with wave-uniform addresses, so they are served by the
// scalar cache / broadcast from LDS, never per-lane gathers.
#pragma once
#include <cstddef>
#include <cstdint>

#if !defined(__HIPCC__) && !defined(__host__)
#define __host__
#define __device__
#endif

namespace sapamd {

enum : int32_t { KIND_REVOLUTE = 0, KIND_PRISMATIC = 1, KIND_CONTINUOUS = 2 };

template <class T> struct alignas(16) JointRec {
  // Joint placement composed with the joint motion, per joint kind:
  //   revolute/continuous: R_local(q) = A + sin(q) Bq + (1 - cos(q)) Cq, p_local = p
  //   prismatic:           R_local    = A,                             p_local = p + ap q
  // with A = R_place, Bq = R_place [a]x, Cq = R_place [a]x^2, ap = R_place a.
  T A[9], Bq[9], Cq[9];
  T p[3], ap[3];
  T axis[3];      // a, the unit joint axis in the joint frame
  T m, c[3];      // mass and centre of mass (joint frame), fixed children merged
  T Ic[6];        // rotational inertia about the COM: xx xy xz yy yz zz (joint frame)
  T lR[9], lp[3]; // placement of the user link attached to this joint (when link >= 0)
  T fpad[2];      // keeps the integer block 16-byte aligned
  int32_t parent; // parent joint, -1 = universe
  int32_t kind;
  int32_t vuser;  // user (set_joint_order) dof index
  int32_t depth;  // 1 for children of the universe
  int32_t nchild;
  int32_t link;   // the one user link attached to this joint, -1 = none (see links_le1)
  int32_t diagI;  // 1 when Ic is diagonal
  int32_t lident; // 1 when the attached link's frame is the joint frame (lR = I, lp = 0)
  int32_t anc[5]; // 2^k-th ancestor (k = 0..4), -1 past the root
  int32_t subtree;  // joints in the subtree (the subtree is [j, j + subtree) in depth-first order)
  uint32_t ancmask; // bit i set when joint i is this joint or one of its ancestors (joints 0..31; 0 past 31)
  int32_t pad2;
};

template <class T> struct alignas(16) LinkRec {
  T R[9], p[3];   // link frame relative to its joint frame
  T c[3];         // the link's own centre of mass in the link frame
  int32_t joint;  // -1 = universe (the root link)
  int32_t ident;  // 1 when R = I and p = 0
  int32_t pad[2];
};

template <class T> struct ModelTable {
  int32_t nj, nlinks, nv, chain;  // chain = 1 when joint j's parent is j - 1 for all j
  int32_t max_depth;
  int32_t ident_order;  // 1 when joint j drives user dof j for every j
  int32_t links_le1;    // 1 when every joint carries at most one user link and at most one
                        // user link sits on the universe: links can be emitted per joint
  int32_t root_link;    // the user link on the universe (-1 = none) when links_le1
  T gravity[4];
  T root_R[9], root_p[3];  // that link's placement
  T root_q[4];             // root_R as a quaternion (w x y z, Eigen's sign rule)
  // JointRec<T> joints[nj]; LinkRec<T> links[nlinks]; follow at kHeaderBytes
};
constexpr size_t kHeaderBytes = 256;

template <class T> __host__ __device__ inline const JointRec<T> *joints_of(const ModelTable<T> *m) {
  return reinterpret_cast<const JointRec<T> *>(reinterpret_cast<const char *>(m) + kHeaderBytes);
}
template <class T> __host__ __device__ inline const LinkRec<T> *links_of(const ModelTable<T> *m) {
  return reinterpret_cast<const LinkRec<T> *>(reinterpret_cast<const char *>(joints_of(m)) +
                                              sizeof(JointRec<T>) * m->nj);
}
static_assert(sizeof(ModelTable<double>) <= kHeaderBytes, "header too large");
static_assert(sizeof(JointRec<float>) % 16 == 0 && sizeof(JointRec<double>) % 16 == 0, "misaligned joint record");

template <class T> inline size_t table_bytes(int nj, int nlinks) {
  return kHeaderBytes + sizeof(JointRec<T>) * nj + sizeof(LinkRec<T>) * nlinks;
}

}  // namespace sapamd
