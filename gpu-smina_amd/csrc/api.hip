/*
 * The C ABI of include/sgpu.h over the gfx950 kernels.  A context owns one
 * device's memory: tables, receptors, maps, model batches and per-call
 * scratch.  Every entry point validates its arguments, launches on the
 * context's stream and synchronises before returning (as the reference's
 * run_* functions do, src/gpu/gpu_kernels_impl.cuh:312, :549, :996).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <array>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sgpu.h"
#include "chain.h"
#include "launch.h"

using namespace sgpu;

namespace
{

struct Fail
{
	sgpu_status code;
	std::string msg;
};

void hip_check(hipError_t e, const char* what)
{
	if (e == hipErrorOutOfMemory)
		throw Fail{SGPU_E_OOM, std::string(what) + ": out of device memory"};
	if (e != hipSuccess)
		throw Fail{SGPU_E_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}
void require(bool ok, const std::string& msg)
{
	if (!ok)
		throw Fail{SGPU_E_INVALID, msg};
}

template<typename T> T Real_host_max();
template<> float Real_host_max<float>() { return FLT_MAX; }
template<> double Real_host_max<double>() { return DBL_MAX; }

// device buffer owned by the context
// The stream of the API call in progress on this thread.  Device buffers
// are allocated and freed in its order: stream-ordered allocation does not
// synchronize the device (hipFree does), so a context's calls never wait for
// another context's kernels.
thread_local hipStream_t t_stream = nullptr;

struct DBuf
{
	void* p = nullptr;
	size_t n = 0;
	hipStream_t st = nullptr; // the stream its allocation and free are ordered on
	DBuf() = default;
	DBuf(const DBuf&) = delete;
	DBuf& operator=(const DBuf&) = delete;
	~DBuf() { reset(); }
	void reset()
	{
		if (p)
			(void)hipFreeAsync(p, st);
		p = nullptr;
		n = 0;
	}
	void alloc(size_t bytes)
	{
		reset();
		st = t_stream;
		n = bytes;
		if (bytes)
			hip_check(hipMallocAsync(&p, bytes, st), "hipMallocAsync");
	}
	template<typename X> X* as() const { return static_cast<X*>(p); }
};
template<typename X>
void upload(DBuf& b, const std::vector<X>& v, hipStream_t st)
{
	b.alloc(std::max<size_t>(1, v.size()) * sizeof(X));
	if (!v.empty())
		hip_check(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice, st), "upload");
}

struct Tables
{
	int n = 0;
	double factor = 0, cutoff_sqr = 0;
	DBuf smooth32, smooth64, fast;
	// charge-dependent terms: the other three components
	int ncomp = 1;
	DBuf qsmooth32, qsmooth64, qfast;
	bool ok = false;
	// precalculate_splines (kind 1)
	int kind = 0;
	DBuf spl, spl_cut;
	double fraction = 0;
	int nseg = 0;
	// free every buffer (their frees are ordered on the context's stream)
	void release()
	{
		for (DBuf* b : {&smooth32, &smooth64, &fast, &qsmooth32, &qsmooth64, &qfast, &spl, &spl_cut})
			b->reset();
	}
};

struct Cells
{
	int off[3], dim[3];
	double begin[3], end[3];
	int enabled[3];
	DBuf start, atoms;
};

struct Receptor
{
	int n = 0;
	std::vector<double> xyz;
	std::vector<int> type;
	std::vector<double> charge;
	DBuf xyz64, xyz32, typ, q64, q32;
	// direct-receptor cells per box (full 3 A bricks)
	std::map<std::vector<double>, std::unique_ptr<Cells>> cells;
};

struct Grid
{
	int receptor = -1;
	int precision = 0;
	sgpu_grid_dims gd;
	int dim[3];
	double init[3], factor[3], factor_inv[3], dim_fl_minus_1[3];
	std::vector<int> types;
	int slot_of[NUM_ATOM_TYPES];
	DBuf points; // FP64 point maps
	DBuf cells;  // corner-packed, precision
	// charge maps (four-component tables), same layouts; empty otherwise
	bool charged = false;
	DBuf qpoints, qcells;
	DBuf slot_dev;
};

struct HostBatch
{
	std::vector<ModelHdr> hdr;
	std::vector<int> iblob, pair_a, pair_b, pair_tri;
	std::vector<double> tblob;
	WavePlan plan{};
};

struct Batch
{
	int precision = 0;
	int n = 0;
	HostBatch h;
	DBuf hdr, iblob, tblob, pair_a, pair_b, pair_tri;
	// FP64 copy for the final rescoring when the batch searches in FP32
	std::unique_ptr<Batch> fp64;
};

// Scratch of the allocation-free entry point (sgpu_eval_confs_noalloc): one
// array per GPU, held until the process ends.  Host code never uses this
// name as an address: each device's address comes from hipGetSymbolAddress
// (noalloc_scratch_of), taken once per device with that device current.
static __device__ __align__(256) unsigned char scratch[SGPU_NOALLOC_SCRATCH_BYTES];

std::mutex noalloc_mu;
std::vector<unsigned char*> noalloc_ptr; // per device, null until taken

// the device address of `scratch` on `device` (current), taken on first use
hipError_t noalloc_scratch_of(int device, unsigned char** out)
{
	std::lock_guard<std::mutex> lock(noalloc_mu);
	if (device >= int(noalloc_ptr.size()))
		noalloc_ptr.resize(size_t(device) + 1, nullptr);
	if (!noalloc_ptr[size_t(device)])
	{
		void* ptr = nullptr;
		const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (e != hipSuccess)
			return e;
		noalloc_ptr[size_t(device)] = static_cast<unsigned char*>(ptr);
	}
	*out = noalloc_ptr[size_t(device)];
	return hipSuccess;
}

} // namespace

struct sgpu_ctx
{
	int device = 0;
	hipStream_t stream = nullptr;
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	// size classes of one Monte Carlo launch run concurrently: the heaviest
	// on `stream`, the others on these (one hardware queue per stream)
	static constexpr int NCLASS = 4;
	hipStream_t class_stream[NCLASS - 1] = {};
	hipEvent_t class_done[NCLASS - 1] = {};
	// recorded on `stream` at the start of every part of a launch: the side
	// streams wait on it, so they see that part's step limits
	hipEvent_t part_start = nullptr;
	std::mutex mu;
	std::string err;
	Tables tab;
	std::map<int, std::unique_ptr<Receptor>> receptors;
	std::map<int, std::shared_ptr<Grid>> grids; // shared between contexts by sgpu_share_grid
	std::map<int, std::unique_ptr<Batch>> batches;
	int next_id = 1;
	std::map<std::string, double> last_ms;
	DBuf scratch; // per-call scratch
	DBuf scratch2;
	DBuf err_flag;
	int progress_chunks = 1;
	sgpu_progress_fn progress_fn = nullptr;
	void* progress_user = nullptr;
	// this device's copy of the allocation-free entry point's scratch
	unsigned char* noalloc_scratch = nullptr;
	// Monte Carlo launch plans per size class: chains per workgroup and the
	// persistent form, as the autotuner measured them (keyed by precision,
	// kernel variant and class), and a one-call override the autotuner's
	// probes set
	struct ClassPlan
	{
		int waves = 0;
		bool persist = false;
	};
	std::map<std::array<int, 3>, ClassPlan> tuned;
	ClassPlan plan_override[NCLASS];
	bool override_on[NCLASS] = {};
	// the user grid (--user_grid): its geometry and corner-packed values
	struct UserGrid
	{
		bool on = false;
		int dim[3] = {0, 0, 0};
		double init[3], factor[3], finv[3], dm1[3];
		double map_slope = 1e6;
		DBuf cells32, cells64;
	} ug;
};

namespace
{

// the LDS plan that fits every listed model of a packed batch
WavePlan plan_for(const HostBatch& B, const std::vector<int>& models)
{
	int nmax = 0, fmax = 0, cmax = 0, gmax = 0, pmax = 0, imax = 0, tmax = 0, amax = 0;
	for (int mi : models)
	{
		const ModelHdr& h = B.hdr[size_t(mi)];
		nmax = std::max(nmax, h.natoms);
		cmax = std::max(cmax, h.conf_size);
		amax = std::max(amax, h.natoms + h.ninflex);
		fmax = std::max(fmax, h.nframes);
		gmax = std::max(gmax, h.change_size);
		pmax = std::max(pmax, h.npairs);
		imax = std::max(imax, h.iblob_len);
		tmax = std::max(tmax, h.tblob_len);
	}
	// tight sizes (multiples of 4): LDS decides how many chains share a CU
	WavePlan p{};
	p.nmax = std::max(4, (nmax + 3) / 4 * 4);
	p.fmax = fmax;
	p.cmax = std::max(8, (cmax + 3) / 4 * 4);
	p.gmax = std::max(8, (gmax + 3) / 4 * 4);
	p.pmax = std::max(1, pmax);
	p.hmax = std::max(1, gmax * gmax);
	p.imax = imax;
	p.tmax = tmax;
	p.amax = amax;
	p.bytes = 0;
	return p;
}

// flat models -> per-model int and real blobs with packed neighbour codes
void pack_batch(const sgpu_flat_model* ms, int n, HostBatch& B)
{
	for (int mi = 0; mi < n; mi++)
	{
		const sgpu_flat_model& m = ms[mi];
		const int na = m.natoms + m.ninflex;
		require(m.natoms > 0 && na <= MAX_ATOMS, "model atoms out of range (1.." + std::to_string(MAX_ATOMS) + ")");
		require(m.nframes > 0 && m.nframes <= MAX_FRAMES, "model frames out of range");
		require(m.conf_size <= MAX_CONF && m.change_size <= MAX_CHANGE, "model conf/change size over 64");
		require(m.ntrees > 0 && m.ntrees <= MAX_TREES && m.nligands <= m.ntrees, "model trees out of range");
		require(m.npairs >= 0 && m.npairs <= MAX_PAIRS && m.nother >= 0 && m.nother <= m.npairs, "model pairs out of range");
		for (int i = 0; i < na; i++)
			require(m.atom_type[i] >= 0 && m.atom_type[i] < NUM_ATOM_TYPES, "atom type out of range");
		ModelHdr h{};
		h.natoms = m.natoms;
		h.ninflex = m.ninflex;
		h.nframes = m.nframes;
		h.ntrees = m.ntrees;
		h.nligands = m.nligands;
		h.npairs = m.npairs;
		h.nother = m.nother;
		h.conf_size = m.conf_size;
		h.change_size = m.change_size;
		h.num_tors = m.num_tors;
		std::vector<int> I;
		std::vector<double> R;
		auto mark = [&](int& off) { off = int(I.size()); };
		// atoms
		mark(h.o_type);
		I.insert(I.end(), m.atom_type, m.atom_type + na);
		mark(h.o_afr);
		for (int i = 0; i < m.natoms; i++)
		{
			require(m.atom_frame[i] >= 0 && m.atom_frame[i] < m.nframes, "atom frame out of range");
			I.push_back(m.atom_frame[i]);
		}
		// neighbour lists in evaluation (pair) order
		std::vector<std::vector<int>> lists(m.natoms);
		for (int p = 0; p < m.npairs; p++)
		{
			const int a = m.pair_a[p], b = m.pair_b[p];
			require(a >= 0 && a < m.natoms && b > a && b < na, "pair atoms out of range");
			const int tri = m.pair_tri[p];
			require(tri >= 0 && tri < SGPU_NUM_TYPE_PAIRS, "pair type index out of range");
			require(tri == tri_index_permissive(m.atom_type[a], m.atom_type[b]), "pair table index disagrees with types");
			const bool other = p < m.nother;
			lists[a].push_back(make_code(b, tri, false, other, p));
			if (b < m.natoms)
				lists[b].push_back(make_code(a, tri, true, other, p));
		}
		mark(h.o_nbr);
		int k = 0;
		for (int i = 0; i < m.natoms; i++)
		{
			I.push_back(k);
			k += int(lists[i].size());
		}
		I.push_back(k);
		mark(h.o_code);
		for (int i = 0; i < m.natoms; i++)
			I.insert(I.end(), lists[i].begin(), lists[i].end());
		if (I.size() % 2)
			I.push_back(0); // pair records are read as 8-byte words
		mark(h.o_prec);
		for (int p = 0; p < m.npairs; p++)
		{
			I.push_back(int(unsigned(3 * m.pair_a[p]) | (unsigned(3 * m.pair_b[p]) << 16)));
			I.push_back(int(unsigned(m.pair_tri[p]) | (unsigned(p < m.nother) << 16)));
		}
		// frames, children, levels
		std::vector<int> depth(m.nframes, 0);
		int nlevels = 0;
		for (int f = 0; f < m.nframes; f++)
		{
			const int p = m.frame_parent[f];
			require(p < f, "frames must be in preorder");
			depth[f] = p < 0 ? 0 : depth[p] + 1;
			nlevels = std::max(nlevels, depth[f] + 1);
			require(m.frame_begin[f] >= 0 && m.frame_end[f] <= m.natoms && m.frame_begin[f] <= m.frame_end[f],
					"frame atom range out of range");
			require(m.frame_kind[f] >= SGPU_FRAME_RIGID_ROOT && m.frame_kind[f] <= SGPU_FRAME_SEGMENT, "frame kind");
			require((m.frame_kind[f] == SGPU_FRAME_SEGMENT) == (p >= 0), "only segments have parents");
		}
		h.nlevels = nlevels;
		mark(h.o_fparent);
		I.insert(I.end(), m.frame_parent, m.frame_parent + m.nframes);
		mark(h.o_fkind);
		I.insert(I.end(), m.frame_kind, m.frame_kind + m.nframes);
		mark(h.o_fbegin);
		I.insert(I.end(), m.frame_begin, m.frame_begin + m.nframes);
		mark(h.o_fend);
		I.insert(I.end(), m.frame_end, m.frame_end + m.nframes);
		mark(h.o_fconf);
		I.insert(I.end(), m.frame_conf, m.frame_conf + m.nframes);
		mark(h.o_fchange);
		I.insert(I.end(), m.frame_change, m.frame_change + m.nframes);
		// end of each frame's subtree atoms: a subtree is a run of frames in
		// preorder whose atom ranges follow each other
		std::vector<int> send(m.frame_end, m.frame_end + m.nframes);
		for (int f = m.nframes - 1; f > 0; f--)
			if (m.frame_parent[f] >= 0)
				send[size_t(m.frame_parent[f])] = std::max(send[size_t(m.frame_parent[f])], send[size_t(f)]);
		for (int f = 0; f + 1 < m.nframes; f++)
			require(m.frame_end[f] == m.frame_begin[f + 1], "frame atom ranges must follow preorder");
		mark(h.o_fsend);
		I.insert(I.end(), send.begin(), send.end());
		// each frame's path from its tree root (frame indices as bytes, root
		// first, the frame itself last) for the ancestor walk of set_conf
		mark(h.o_fdepth);
		h.path_ok = 1;
		for (int f = 0; f < m.nframes; f++)
		{
			I.push_back(depth[size_t(f)] + 1);
			if (depth[size_t(f)] + 1 > MAX_PATH)
				h.path_ok = 0;
		}
		mark(h.o_fpath);
		for (int f = 0; f < m.nframes; f++)
		{
			std::vector<int> path;
			for (int a = f; a >= 0; a = m.frame_parent[a])
				path.push_back(a);
			std::reverse(path.begin(), path.end());
			path.resize(MAX_PATH, 0);
			for (int w = 0; w < MAX_PATH / 4; w++)
				I.push_back(int(unsigned(path[size_t(4 * w)]) | (unsigned(path[size_t(4 * w + 1)]) << 8) |
						(unsigned(path[size_t(4 * w + 2)]) << 16) | (unsigned(path[size_t(4 * w + 3)]) << 24)));
		}
		std::vector<int> cstart, ccount, clist;
		for (int f = 0; f < m.nframes; f++)
		{
			cstart.push_back(int(clist.size()));
			int cnt = 0;
			for (int c = f + 1; c < m.nframes; c++)
				if (m.frame_parent[c] == f)
				{
					clist.push_back(c);
					++cnt;
				}
			ccount.push_back(cnt);
		}
		mark(h.o_cstart);
		I.insert(I.end(), cstart.begin(), cstart.end());
		mark(h.o_ccount);
		I.insert(I.end(), ccount.begin(), ccount.end());
		mark(h.o_clist);
		I.insert(I.end(), clist.begin(), clist.end());
		mark(h.o_lstart);
		std::vector<int> lframes;
		for (int L = 0; L < nlevels; L++)
		{
			I.push_back(int(lframes.size()));
			for (int f = 0; f < m.nframes; f++)
				if (depth[f] == L)
					lframes.push_back(f);
		}
		I.push_back(int(lframes.size()));
		mark(h.o_lframes);
		I.insert(I.end(), lframes.begin(), lframes.end());
		// heavy atoms
		mark(h.o_heavy);
		int nh = 0;
		for (int i = 0; i < m.natoms; i++)
			if (!is_hydrogen_type(m.atom_type[i]))
			{
				I.push_back(i);
				++nh;
			}
		h.nheavy = nh;
		// trees
		std::vector<int> troot(m.ntrees, -1), tab(m.ntrees, m.natoms), tae(m.ntrees, 0), tn(m.ntrees, 0);
		for (int f = 0; f < m.nframes; f++)
		{
			const int t = m.frame_tree[f];
			require(t >= 0 && t < m.ntrees, "frame tree out of range");
			if (m.frame_parent[f] < 0)
				troot[t] = f;
			tab[t] = std::min(tab[t], m.frame_begin[f]);
			tae[t] = std::max(tae[t], m.frame_end[f]);
			if (m.frame_kind[f] != SGPU_FRAME_RIGID_ROOT)
				++tn[t];
		}
		std::vector<int> tconf(m.ntrees), tchg(m.ntrees), tlig(m.ntrees), tpb(m.ntrees), tpe(m.ntrees);
		for (int t = 0; t < m.ntrees; t++)
		{
			require(troot[t] >= 0, "tree without a root frame");
			tconf[t] = m.frame_conf[troot[t]];
			tchg[t] = m.frame_change[troot[t]];
			tlig[t] = t < m.nligands ? 1 : 0;
			// a ligand's internal pairs: contiguous after the other pairs
			int pb = m.npairs, pe = m.nother;
			if (tlig[t])
				for (int p = m.nother; p < m.npairs; p++)
					if (m.pair_a[p] >= tab[t] && m.pair_a[p] < tae[t])
					{
						pb = std::min(pb, p);
						pe = std::max(pe, p + 1);
					}
			if (pb > pe)
				pb = pe;
			tpb[t] = pb;
			tpe[t] = pe;
		}
		mark(h.o_troot);
		I.insert(I.end(), troot.begin(), troot.end());
		mark(h.o_tconf);
		I.insert(I.end(), tconf.begin(), tconf.end());
		mark(h.o_tchange);
		I.insert(I.end(), tchg.begin(), tchg.end());
		mark(h.o_tlig);
		I.insert(I.end(), tlig.begin(), tlig.end());
		mark(h.o_tntors);
		I.insert(I.end(), tn.begin(), tn.end());
		mark(h.o_tab);
		I.insert(I.end(), tab.begin(), tab.end());
		mark(h.o_tae);
		I.insert(I.end(), tae.begin(), tae.end());
		mark(h.o_tpb);
		I.insert(I.end(), tpb.begin(), tpb.end());
		mark(h.o_tpe);
		I.insert(I.end(), tpe.begin(), tpe.end());
		// conf / change maps
		std::vector<int> ck(m.conf_size, -1), cc(m.conf_size, 0), gk(m.change_size, -1), gc(m.change_size, 0);
		for (int f = 0; f < m.nframes; f++)
		{
			const int ci = m.frame_conf[f], gi = m.frame_change[f];
			if (m.frame_kind[f] == SGPU_FRAME_RIGID_ROOT)
			{
				require(ci >= 0 && ci + 7 <= m.conf_size && gi >= 0 && gi + 6 <= m.change_size,
						"root conf index out of range");
				for (int d = 0; d < 3; d++)
				{
					ck[ci + d] = CK_POSITION;
					cc[ci + d] = gi + d;
					gk[gi + d] = GK_POSITION;
					gc[gi + d] = ci + d;
				}
				for (int d = 0; d < 4; d++)
				{
					ck[ci + 3 + d] = CK_QUAT0 + d;
					cc[ci + 3 + d] = gi + 3;
				}
				for (int d = 0; d < 3; d++)
				{
					gk[gi + 3 + d] = GK_ORIENT0 + d;
					gc[gi + 3 + d] = ci + 3;
				}
			}
			else
			{
				require(ci >= 0 && ci < m.conf_size && gi >= 0 && gi < m.change_size, "torsion index out of range");
				ck[ci] = CK_TORSION;
				cc[ci] = gi;
				gk[gi] = GK_TORSION;
				gc[gi] = ci;
			}
		}
		for (int v : ck)
			require(v >= 0, "conf element without a frame");
		for (int v : gk)
			require(v >= 0, "change element without a frame");
		mark(h.o_maps);
		I.insert(I.end(), ck.begin(), ck.end());
		I.insert(I.end(), cc.begin(), cc.end());
		I.insert(I.end(), gk.begin(), gk.end());
		I.insert(I.end(), gc.begin(), gc.end());
		// reals
		h.o_xyz0 = int(R.size());
		R.insert(R.end(), m.atom_init, m.atom_init + 3 * na);
		h.o_local = int(R.size());
		R.insert(R.end(), m.atom_local, m.atom_local + 3 * m.natoms);
		h.o_forig = int(R.size());
		R.insert(R.end(), m.frame_origin, m.frame_origin + 3 * m.nframes);
		h.o_faxis = int(R.size());
		R.insert(R.end(), m.frame_axis, m.frame_axis + 3 * m.nframes);
		h.o_charge = int(R.size());
		if (m.atom_charge)
			R.insert(R.end(), m.atom_charge, m.atom_charge + na);
		else
			R.insert(R.end(), size_t(na), 0.0);
		// append
		h.iblob_off = int(B.iblob.size());
		h.iblob_len = int(I.size());
		h.tblob_off = int(B.tblob.size());
		h.tblob_len = int(R.size());
		h.pair_off = int(B.pair_a.size());
		B.iblob.insert(B.iblob.end(), I.begin(), I.end());
		B.tblob.insert(B.tblob.end(), R.begin(), R.end());
		B.pair_a.insert(B.pair_a.end(), m.pair_a, m.pair_a + m.npairs);
		B.pair_b.insert(B.pair_b.end(), m.pair_b, m.pair_b + m.npairs);
		B.pair_tri.insert(B.pair_tri.end(), m.pair_tri, m.pair_tri + m.npairs);
		B.hdr.push_back(h);
	}
	std::vector<int> all(static_cast<size_t>(n));
	for (int mi = 0; mi < n; mi++)
		all[size_t(mi)] = mi;
	B.plan = plan_for(B, all);
}

template<typename T>
std::vector<T> cast(const std::vector<double>& v)
{
	return std::vector<T>(v.begin(), v.end());
}

void upload_batch(Batch& b, hipStream_t st)
{
	const HostBatch& h = b.h;
	b.hdr.alloc(h.hdr.size() * sizeof(ModelHdr));
	hip_check(hipMemcpyAsync(b.hdr.p, h.hdr.data(), h.hdr.size() * sizeof(ModelHdr), hipMemcpyHostToDevice, st), "hdr");
	upload(b.iblob, h.iblob, st);
	upload(b.pair_a, h.pair_a, st);
	upload(b.pair_b, h.pair_b, st);
	upload(b.pair_tri, h.pair_tri, st);
	if (b.precision == SGPU_FP64)
		upload(b.tblob, h.tblob, st);
	else
		upload(b.tblob, cast<float>(h.tblob), st);
	hip_check(hipStreamSynchronize(st), "upload models");
}

template<typename T>
BatchPtrs<T> ptrs(const Batch& b)
{
	BatchPtrs<T> p;
	p.hdr = b.hdr.as<ModelHdr>();
	p.iblob = b.iblob.as<int>();
	p.tblob = b.tblob.as<T>();
	p.pair_a = b.pair_a.as<int>();
	p.pair_b = b.pair_b.as<int>();
	p.pair_tri = b.pair_tri.as<int>();
	return p;
}

int plan_bytes(const WavePlan& p, int prec, int waves)
{
	return prec == SGPU_FP64 ? block_lds_bytes<double>(p, waves)
	       : prec == SGPU_MIXED ? block_lds_bytes<float, double>(p, waves)
	                            : block_lds_bytes<float>(p, waves);
}
// the precision of a batch's model evaluation (and so of its maps)
int eval_precision(int prec) { return prec == SGPU_FP64 ? SGPU_FP64 : SGPU_FP32; }
constexpr int LDS_LIMIT = 160 * 1024;
constexpr int LDS_STATIC = 64; // the persistent Monte Carlo kernel's own LDS (its work item)
constexpr double TUNE_MARGIN = 0.03; // a tuned plan must beat the default by this fraction
// the speculative-chain kernel's own LDS (its request and result exchange)
template<typename S>
constexpr int spec_static_bytes()
{
	return int(sizeof(S)) * (SPEC_REQ + SPEC_MAX * SPEC_RS) + 64;
}

template<typename T>
TablePtrs<T> table_ptrs(const sgpu_ctx* c)
{
	TablePtrs<T> t;
	t.smooth = (sizeof(T) == 8 ? c->tab.smooth64 : c->tab.smooth32).template as<T>();
	t.fast = c->tab.fast.as<double>();
	t.ncomp = c->tab.ncomp;
	t.qsmooth = (sizeof(T) == 8 ? c->tab.qsmooth64 : c->tab.qsmooth32).template as<T>();
	t.qfast = c->tab.qfast.as<double>();
	t.n = c->tab.n;
	t.factor = c->tab.factor;
	t.cutoff_sqr = c->tab.cutoff_sqr;
	t.kind = c->tab.kind;
	t.spl = c->tab.spl.as<double>();
	t.spl_cut = c->tab.spl_cut.as<double>();
	t.fraction = c->tab.fraction;
	t.nseg = c->tab.nseg;
	return t;
}


template<typename T>
GridPtrs<T> grid_ptrs(const Grid& g, double slope)
{
	GridPtrs<T> p;
	p.cells = g.cells.as<T>();
	p.qcells = g.charged ? g.qcells.as<T>() : nullptr;
	p.slot_of = g.slot_dev.as<int>();
	for (int d = 0; d < 3; d++)
	{
		p.dim[d] = g.dim[d];
		p.init[d] = g.init[d];
		p.factor[d] = g.factor[d];
		p.factor_inv[d] = g.factor_inv[d];
		p.dim_fl_minus_1[d] = g.dim_fl_minus_1[d];
	}
	p.ncells = (g.dim[0] - 1) * (g.dim[1] - 1) * (g.dim[2] - 1);
	p.slope = slope;
	return p;
}

// grid::init (src/lib/grid.cpp:60-75)
void grid_geometry(const sgpu_grid_dims& gd, int* dim, double* init, double* factor, double* finv, double* dm1)
{
	for (int i = 0; i < 3; i++)
	{
		dim[i] = gd.n[i] + 1;
		init[i] = gd.begin[i];
		const double range = gd.end[i] - gd.begin[i];
		dm1[i] = dim[i] - 1.0;
		factor[i] = dm1[i] / range;
		finv[i] = 1 / factor[i];
	}
}

// the user grid as a one-map GridPtrs (slot 0); has_ug 0 without one
template<typename T>
void attach_user_grid(const sgpu_ctx* c, EvalCtx<T>& ec)
{
	ec.has_ug = c->ug.on ? 1 : 0;
	if (!c->ug.on)
		return;
	GridPtrs<T>& p = ec.ug;
	p.cells = (sizeof(T) == 8 ? c->ug.cells64 : c->ug.cells32).template as<T>();
	p.qcells = nullptr;
	p.slot_of = nullptr;
	for (int d = 0; d < 3; d++)
	{
		p.dim[d] = c->ug.dim[d];
		p.init[d] = c->ug.init[d];
		p.factor[d] = c->ug.factor[d];
		p.factor_inv[d] = c->ug.finv[d];
		p.dim_fl_minus_1[d] = c->ug.dm1[d];
	}
	p.ncells = (p.dim[0] - 1) * (p.dim[1] - 1) * (p.dim[2] - 1);
	p.slope = 0;
}

void check_gd(const sgpu_grid_dims* gd)
{
	require(gd != nullptr, "grid dims missing");
	for (int i = 0; i < 3; i++)
	{
		require(gd->n[i] >= 1 && gd->n[i] <= 4096, "grid intervals out of range");
		require(gd->end[i] > gd->begin[i], "grid box end must exceed begin");
	}
}

// cells of the direct back end over box gd: full 3 A bricks
Cells& receptor_cells(sgpu_ctx* c, Receptor& r, const sgpu_grid_dims& gd)
{
	std::vector<double> key;
	for (int i = 0; i < 3; i++)
	{
		key.push_back(gd.begin[i]);
		key.push_back(gd.end[i]);
		key.push_back(gd.n[i]);
	}
	auto it = r.cells.find(key);
	if (it != r.cells.end())
		return *it->second;
	std::unique_ptr<Cells> cl(new Cells);
	GridBuildArgs a{};
	a.rec_xyz = r.xyz64.as<double>();
	a.rec_type = r.typ.as<int>();
	a.rec_n = r.n;
	a.cutoff_sqr = c->tab.cutoff_sqr;
	std::vector<double> lo, hi;
	for (int d = 0; d < 3; d++)
	{
		cl->off[d] = int(std::floor(gd.begin[d] / 3.0));
		cl->dim[d] = int(std::ceil(gd.end[d] / 3.0)) - cl->off[d] + 1;
		cl->begin[d] = gd.begin[d];
		cl->end[d] = gd.end[d];
		cl->enabled[d] = gd.n[d] > 0;
		a.cell_off[d] = cl->off[d];
		a.cell_dim[d] = cl->dim[d];
		a.box_lo[d] = gd.begin[d];
		a.box_hi[d] = gd.end[d];
	}
	for (int d = 0; d < 3; d++)
		for (int i = 0; i < cl->dim[d]; i++)
		{
			lo.push_back((cl->off[d] + i) * 3.0);
			hi.push_back((cl->off[d] + i + 1) * 3.0);
		}
	DBuf dlo, dhi;
	upload(dlo, lo, c->stream);
	upload(dhi, hi, c->stream);
	a.cell_lo = dlo.as<double>();
	a.cell_hi = dhi.as<double>();
	const int ncells = cl->dim[0] * cl->dim[1] * cl->dim[2];
	cl->start.alloc(size_t(ncells + 1) * sizeof(int));
	a.cell_start = cl->start.as<int>();
	hip_check(launch_cell_count(a, c->stream), "cell count");
	int total = 0;
	hip_check(hipMemcpyAsync(&total, a.cell_start + ncells, sizeof(int), hipMemcpyDeviceToHost, c->stream), "cells");
	hip_check(hipStreamSynchronize(c->stream), "cells");
	cl->atoms.alloc(size_t(std::max(1, total)) * sizeof(int));
	a.cell_atoms = cl->atoms.as<int>();
	hip_check(launch_cell_fill(a, c->stream), "cell fill");
	hip_check(hipStreamSynchronize(c->stream), "cells");
	Cells& ref = *cl;
	r.cells[key] = std::move(cl);
	return ref;
}

template<typename T>
ReceptorPtrs<T> receptor_ptrs(const Receptor& r, const Cells& cl)
{
	ReceptorPtrs<T> p;
	p.xyz = (sizeof(T) == 8 ? r.xyz64 : r.xyz32).template as<T>();
	p.type = r.typ.as<int>();
	p.charge = (sizeof(T) == 8 ? r.q64 : r.q32).template as<T>();
	p.natoms = r.n;
	p.cell_start = cl.start.as<int>();
	p.cell_atoms = cl.atoms.as<int>();
	for (int d = 0; d < 3; d++)
	{
		p.cell_off[d] = cl.off[d];
		p.cell_dim[d] = cl.dim[d];
		p.begin[d] = cl.begin[d];
		p.end[d] = cl.end[d];
		p.enabled[d] = cl.enabled[d];
	}
	return p;
}

template<typename F>
sgpu_status guarded(sgpu_ctx* c, F&& f)
{
	if (!c)
		return SGPU_E_INVALID;
	std::lock_guard<std::mutex> lock(c->mu);
	struct StreamScope
	{
		hipStream_t saved;
		explicit StreamScope(hipStream_t s) : saved(t_stream) { t_stream = s; }
		~StreamScope() { t_stream = saved; }
	} scope(c->stream);
	try
	{
		hip_check(hipSetDevice(c->device), "hipSetDevice");
		f();
		return SGPU_OK;
	}
	catch (Fail& e)
	{
		c->err = e.msg;
		return e.code;
	}
	catch (std::exception& e)
	{
		c->err = e.what();
		return SGPU_E_INVALID;
	}
}

struct Timer
{
	sgpu_ctx* c;
	const char* kind;
	Timer(sgpu_ctx* c_, const char* k) : c(c_), kind(k) { hip_check(hipEventRecord(c->ev0, c->stream), "event"); }
	void stop()
	{
		hip_check(hipEventRecord(c->ev1, c->stream), "event");
		hip_check(hipEventSynchronize(c->ev1), "event");
		float ms = 0;
		hip_check(hipEventElapsedTime(&ms, c->ev0, c->ev1), "event");
		c->last_ms[kind] = ms;
	}
};

Batch& find_batch(sgpu_ctx* c, int id)
{
	auto it = c->batches.find(id);
	if (it == c->batches.end())
		throw Fail{SGPU_E_NOT_FOUND, "unknown model batch id"};
	return *it->second;
}
Receptor& find_receptor(sgpu_ctx* c, int id)
{
	auto it = c->receptors.find(id);
	if (it == c->receptors.end())
		throw Fail{SGPU_E_NOT_FOUND, "unknown receptor id"};
	return *it->second;
}
Grid& find_grid(sgpu_ctx* c, int id)
{
	auto it = c->grids.find(id);
	if (it == c->grids.end())
		throw Fail{SGPU_E_NOT_FOUND, "unknown grid id"};
	return *it->second;
}

void ensure_scratch(DBuf& b, size_t bytes)
{
	if (b.n < bytes)
		b.alloc(bytes);
}

// every heavy movable atom of the listed models must have a map
void check_maps(const Batch& b, const Grid& g, const int* models, int n)
{
	for (int i = 0; i < n; i++)
	{
		const ModelHdr& h = b.h.hdr[models[i]];
		for (int a = 0; a < h.natoms; a++)
		{
			const int t = b.h.iblob[h.iblob_off + h.o_type + a];
			require(is_hydrogen_type(t) || g.slot_of[t] >= 0, "a ligand atom type has no map");
		}
	}
}

} // namespace

extern "C" {

sgpu_status sgpu_create(int device, sgpu_ctx** out)
{
	if (!out)
		return SGPU_E_INVALID;
	*out = nullptr;
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
		return SGPU_E_NO_DEVICE;
	if (device < 0 || device >= n)
		return SGPU_E_NO_DEVICE;
	std::unique_ptr<sgpu_ctx> c(new sgpu_ctx);
	c->device = device;
	if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
			hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
		return SGPU_E_HIP;
	for (int k = 0; k < sgpu_ctx::NCLASS - 1; k++)
		if (hipStreamCreateWithFlags(&c->class_stream[k], hipStreamNonBlocking) != hipSuccess ||
				hipEventCreateWithFlags(&c->class_done[k], hipEventDisableTiming) != hipSuccess)
			return SGPU_E_HIP;
	if (hipEventCreateWithFlags(&c->part_start, hipEventDisableTiming) != hipSuccess)
		return SGPU_E_HIP;
	// the allocation-free entry point's scratch address on this device (a
	// failure here disables that entry point only: it reports the error)
	if (noalloc_scratch_of(device, &c->noalloc_scratch) != hipSuccess)
	{
		c->noalloc_scratch = nullptr;
		(void)hipGetLastError();
	}
	*out = c.release();
	return SGPU_OK;
}

void sgpu_destroy(sgpu_ctx* c)
{
	if (!c)
		return;
	(void)hipSetDevice(c->device);
	(void)hipStreamSynchronize(c->stream);
	// every buffer's free is ordered on the context's stream: release them
	// all before the stream goes
	c->batches.clear();
	c->grids.clear();
	c->receptors.clear();
	c->tab.release();
	c->scratch.reset();
	c->scratch2.reset();
	c->err_flag.reset();
	c->ug.cells32.reset();
	c->ug.cells64.reset();
	(void)hipStreamSynchronize(c->stream);
	(void)hipEventDestroy(c->ev0);
	(void)hipEventDestroy(c->ev1);
	(void)hipEventDestroy(c->part_start);
	for (int k = 0; k < sgpu_ctx::NCLASS - 1; k++)
	{
		(void)hipStreamSynchronize(c->class_stream[k]);
		(void)hipEventDestroy(c->class_done[k]);
		(void)hipStreamDestroy(c->class_stream[k]);
	}
	(void)hipStreamDestroy(c->stream);
	delete c;
}

const char* sgpu_last_error(const sgpu_ctx* c) { return c ? c->err.c_str() : "no context"; }

sgpu_status sgpu_device_info(sgpu_ctx* c, char* name, int name_len, int* cus)
{
	return guarded(c, [&]() {
		hipDeviceProp_t p;
		hip_check(hipGetDeviceProperties(&p, c->device), "hipGetDeviceProperties");
		if (name && name_len > 0)
			std::snprintf(name, size_t(name_len), "%s (%s)", p.name, p.gcnArchName);
		if (cus)
			*cus = p.multiProcessorCount;
	});
}

sgpu_status sgpu_device_memory(sgpu_ctx* c, size_t* free_bytes, size_t* total_bytes)
{
	return guarded(c, [&]() {
		require(free_bytes && total_bytes, "no output");
		hip_check(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
	});
}

sgpu_status sgpu_mc_capacity(sgpu_ctx* c, int precision, int* chains)
{
	return guarded(c, [&]() {
		require(chains && (precision == SGPU_FP32 || precision == SGPU_FP64 || precision == SGPU_MIXED),
				"capacity request");
		hipDeviceProp_t p;
		hip_check(hipGetDeviceProperties(&p, c->device), "hipGetDeviceProperties");
		const int per_cu = precision == SGPU_FP64    ? mc_blocks_per_cu<double, double>(1, 0)
		                   : precision == SGPU_MIXED ? mc_blocks_per_cu<float, double>(1, 0)
		                                             : mc_blocks_per_cu<float, float>(1, 0);
		require(per_cu > 0, "occupancy query failed");
		*chains = per_cu * p.multiProcessorCount;
	});
}

sgpu_status sgpu_upload_splines(sgpu_ctx* c, const sgpu_splines* s)
{
	return guarded(c, [&]() {
		require(s && s->coef && s->cut && s->nseg >= 2 && s->nseg <= 100000 && s->fraction > 0 && s->cutoff_sqr > 0,
				"invalid splines");
		const size_t ncoef = size_t(SGPU_NUM_TYPE_PAIRS) * s->nseg * 5;
		upload(c->tab.spl, std::vector<double>(s->coef, s->coef + ncoef), c->stream);
		upload(c->tab.spl_cut, std::vector<double>(s->cut, s->cut + SGPU_NUM_TYPE_PAIRS), c->stream);
		c->tab.fraction = s->fraction;
		c->tab.nseg = s->nseg;
		c->tab.cutoff_sqr = s->cutoff_sqr;
		c->tab.kind = 1;
		c->tab.ncomp = 1; // one component: charge-dependent terms are refused with splines
		c->tab.ok = true;
		hip_check(hipStreamSynchronize(c->stream), "upload splines");
	});
}

sgpu_status sgpu_upload_tables(sgpu_ctx* c, const sgpu_tables* t)
{
	return guarded(c, [&]() {
		require(t && t->smooth && t->fast && t->n >= 4 && t->factor > 0, "invalid tables");
		require(size_t(t->cutoff_sqr * t->factor) + 1 < size_t(t->n), "table too short for the cutoff");
		const size_t ne = size_t(SGPU_NUM_TYPE_PAIRS) * t->n;
		std::vector<double> sm(t->smooth, t->smooth + 2 * ne), fa(t->fast, t->fast + ne);
		upload(c->tab.smooth64, sm, c->stream);
		upload(c->tab.smooth32, cast<float>(sm), c->stream);
		upload(c->tab.fast, fa, c->stream);
		require(t->ncomp == 1 || t->ncomp == 4, "tables have 1 or 4 components");
		require(t->ncomp == 1 || (t->qsmooth && t->qfast), "charge components missing");
		if (t->ncomp == 4)
		{
			std::vector<double> qs(t->qsmooth, t->qsmooth + 6 * ne), qf(t->qfast, t->qfast + 3 * ne);
			upload(c->tab.qsmooth64, qs, c->stream);
			upload(c->tab.qsmooth32, cast<float>(qs), c->stream);
			upload(c->tab.qfast, qf, c->stream);
		}
		c->tab.ncomp = t->ncomp;
		c->tab.n = t->n;
		c->tab.factor = t->factor;
		c->tab.cutoff_sqr = t->cutoff_sqr;
		c->tab.kind = 0;
		c->tab.ok = true;
		hip_check(hipStreamSynchronize(c->stream), "upload tables");
	});
}

sgpu_status sgpu_set_user_grid(sgpu_ctx* c, const sgpu_user_grid* ug)
{
	return guarded(c, [&]() {
		if (!ug)
		{
			c->ug.on = false;
			return;
		}
		check_gd(&ug->gd);
		require(ug->data != nullptr, "user grid values missing");
		require(ug->map_slope >= 0, "user grid map slope");
		sgpu_ctx::UserGrid& u = c->ug;
		grid_geometry(ug->gd, u.dim, u.init, u.factor, u.finv, u.dm1);
		u.map_slope = ug->map_slope;
		// corner-packed cells, as the receptor maps (x fastest)
		const int d0 = u.dim[0], d1 = u.dim[1], d2 = u.dim[2];
		auto at = [&](int x, int y, int z) { return ug->data[size_t(x) + size_t(d0) * (size_t(y) + size_t(d1) * z)]; };
		std::vector<double> cells(size_t(d0 - 1) * (d1 - 1) * (d2 - 1) * 8);
		size_t k = 0;
		for (int z = 0; z < d2 - 1; z++)
			for (int y = 0; y < d1 - 1; y++)
				for (int x = 0; x < d0 - 1; x++)
				{
					const double v[8] = {at(x, y, z), at(x + 1, y, z), at(x, y + 1, z), at(x + 1, y + 1, z),
							at(x, y, z + 1), at(x + 1, y, z + 1), at(x, y + 1, z + 1), at(x + 1, y + 1, z + 1)};
					for (int j = 0; j < 8; j++)
						cells[k++] = v[j];
				}
		upload(u.cells64, cells, c->stream);
		upload(u.cells32, cast<float>(cells), c->stream);
		hip_check(hipStreamSynchronize(c->stream), "user grid");
		u.on = true;
	});
}

sgpu_status sgpu_upload_receptor(sgpu_ctx* c, const sgpu_receptor* r, int* id)
{
	return guarded(c, [&]() {
		require(r && id && r->natoms > 0 && r->xyz && r->type, "invalid receptor");
		std::unique_ptr<Receptor> R(new Receptor);
		R->n = r->natoms;
		R->xyz.assign(r->xyz, r->xyz + 3 * size_t(r->natoms));
		R->type.assign(r->type, r->type + r->natoms);
		for (int t : R->type)
			require(t >= 0 && t < SGPU_NUM_TYPES, "receptor atom type out of range");
		if (r->charge)
			R->charge.assign(r->charge, r->charge + r->natoms);
		else
			R->charge.assign(size_t(r->natoms), 0.0);
		upload(R->xyz64, R->xyz, c->stream);
		upload(R->xyz32, cast<float>(R->xyz), c->stream);
		upload(R->typ, R->type, c->stream);
		upload(R->q64, R->charge, c->stream);
		upload(R->q32, cast<float>(R->charge), c->stream);
		hip_check(hipStreamSynchronize(c->stream), "upload receptor");
		*id = c->next_id++;
		c->receptors[*id] = std::move(R);
	});
}

sgpu_status sgpu_build_grids(sgpu_ctx* c, int receptor_id, const sgpu_grid_dims* gd, const int* types, int ntypes,
		int precision, int* grid_id)
{
	return guarded(c, [&]() {
		require(c->tab.ok, "tables not uploaded");
		check_gd(gd);
		require(grid_id && types && ntypes > 0 && ntypes <= 16, "1..16 map types");
		require(precision == SGPU_FP32 || precision == SGPU_FP64, "invalid precision");
		Receptor& r = find_receptor(c, receptor_id);
		std::shared_ptr<Grid> g(new Grid);
		g->receptor = receptor_id;
		g->precision = precision;
		g->gd = *gd;
		grid_geometry(*gd, g->dim, g->init, g->factor, g->factor_inv, g->dim_fl_minus_1);
		std::fill(g->slot_of, g->slot_of + NUM_ATOM_TYPES, -1);
		for (int j = 0; j < ntypes; j++)
		{
			require(types[j] >= 0 && types[j] < NUM_ATOM_TYPES, "map type out of range");
			require(g->slot_of[types[j]] < 0, "duplicate map type");
			g->slot_of[types[j]] = j;
			g->types.push_back(types[j]);
		}
		// cells with smina's first-probe bricks, per dimension
		GridBuildArgs a{};
		a.rec_xyz = r.xyz64.as<double>();
		a.rec_type = r.typ.as<int>();
		a.rec_q = r.q64.as<double>();
		a.rec_n = r.n;
		a.fast = c->tab.fast.as<double>();
		a.ncomp = c->tab.kind == 0 ? c->tab.ncomp : 1;
		a.qfast = c->tab.qfast.as<double>();
		a.table_n = c->tab.n;
		a.factor = c->tab.factor;
		a.cutoff_sqr = c->tab.cutoff_sqr;
		a.kind = c->tab.kind;
		a.spl = c->tab.spl.as<double>();
		a.spl_cut = c->tab.spl_cut.as<double>();
		a.fraction = c->tab.fraction;
		a.nseg = c->tab.nseg;
		std::vector<double> lo, hi;
		for (int d = 0; d < 3; d++)
		{
			a.cell_off[d] = int(std::floor(gd->begin[d] / 3.0));
			a.cell_dim[d] = int(std::ceil(gd->end[d] / 3.0)) - a.cell_off[d] + 1;
			a.box_lo[d] = gd->begin[d];
			a.box_hi[d] = gd->end[d];
			a.dim[d] = g->dim[d];
			a.init[d] = g->init[d];
			a.factor_inv[d] = g->factor_inv[d];
		}
		for (int d = 0; d < 3; d++)
		{
			std::vector<double> l(a.cell_dim[d], 0), h(a.cell_dim[d], 0);
			std::vector<char> seen(a.cell_dim[d], 0);
			for (int i = 0; i <= gd->n[d]; i++)
			{
				const double cval = g->init[d] + g->factor_inv[d] * i;
				const int cell = int(std::floor(cval / 3.0)) - a.cell_off[d];
				require(cell >= 0 && cell < a.cell_dim[d], "probe outside the cell range");
				if (!seen[cell])
				{
					seen[cell] = 1;
					l[cell] = std::floor(cval / 3.0) * 3.0;
					h[cell] = std::ceil(cval / 3.0) * 3.0;
				}
			}
			lo.insert(lo.end(), l.begin(), l.end());
			hi.insert(hi.end(), h.begin(), h.end());
		}
		DBuf dlo, dhi, dstart, datoms, dtypes;
		upload(dlo, lo, c->stream);
		upload(dhi, hi, c->stream);
		upload(dtypes, g->types, c->stream);
		a.cell_lo = dlo.as<double>();
		a.cell_hi = dhi.as<double>();
		a.types = dtypes.as<int>();
		a.ntypes = ntypes;
		EvalCtx<double> ugc{};
		attach_user_grid(c, ugc);
		a.has_ug = ugc.has_ug;
		a.ug = ugc.ug;
		a.ug_slope = c->ug.map_slope;
		const int ncells = a.cell_dim[0] * a.cell_dim[1] * a.cell_dim[2];
		dstart.alloc(size_t(ncells + 1) * sizeof(int));
		a.cell_start = dstart.as<int>();
		Timer tm(c, "grid_build");
		hip_check(launch_cell_count(a, c->stream), "cell count");
		int total = 0;
		hip_check(hipMemcpyAsync(&total, a.cell_start + ncells, sizeof(int), hipMemcpyDeviceToHost, c->stream), "cells");
		hip_check(hipStreamSynchronize(c->stream), "cells");
		datoms.alloc(size_t(std::max(1, total)) * sizeof(int));
		a.cell_atoms = datoms.as<int>();
		hip_check(launch_cell_fill(a, c->stream), "cell fill");
		const size_t npts = size_t(g->dim[0]) * g->dim[1] * g->dim[2];
		g->points.alloc(npts * ntypes * sizeof(double));
		g->charged = a.ncomp == 4;
		if (g->charged)
			g->qpoints.alloc(npts * ntypes * sizeof(double));
		hip_check(launch_grid_points(a, g->points.as<double>(), g->qpoints.as<double>(), c->stream), "grid points");
		const size_t ncell = size_t(g->dim[0] - 1) * (g->dim[1] - 1) * (g->dim[2] - 1) * ntypes * 8;
		for (int q = 0; q < (g->charged ? 2 : 1); q++)
		{
			DBuf& pts = q ? g->qpoints : g->points;
			DBuf& cells = q ? g->qcells : g->cells;
			if (precision == SGPU_FP64)
			{
				cells.alloc(ncell * sizeof(double));
				hip_check(launch_grid_pack<double>(pts.as<double>(), ntypes, g->dim, cells.as<double>(), c->stream),
						"grid pack");
			}
			else
			{
				cells.alloc(ncell * sizeof(float));
				hip_check(launch_grid_pack<float>(pts.as<double>(), ntypes, g->dim, cells.as<float>(), c->stream),
						"grid pack");
			}
		}
		std::vector<int> slots(g->slot_of, g->slot_of + NUM_ATOM_TYPES);
		upload(g->slot_dev, slots, c->stream);
		tm.stop();
		hip_check(hipStreamSynchronize(c->stream), "build grids");
		*grid_id = c->next_id++;
		c->grids[*grid_id] = std::move(g);
	});
}

sgpu_status sgpu_share_grid(sgpu_ctx* dst, sgpu_ctx* src, int src_grid, int* grid_id)
{
	if (!src || src == dst)
		return SGPU_E_INVALID;
	std::shared_ptr<Grid> g;
	{
		std::lock_guard<std::mutex> lock(src->mu);
		auto it = src->grids.find(src_grid);
		if (it == src->grids.end())
		{
			src->err = "grid not found";
			return SGPU_E_NOT_FOUND;
		}
		g = it->second;
	}
	return guarded(dst, [&]() {
		require(grid_id != nullptr, "no grid id");
		require(src->device == dst->device, "the contexts are on different devices");
		*grid_id = dst->next_id++;
		dst->grids[*grid_id] = g;
	});
}

namespace
{
void download_map(sgpu_ctx* c, int grid_id, int type, bool charge, double* out)
{
	Grid& g = find_grid(c, grid_id);
	require(out && type >= 0 && type < NUM_ATOM_TYPES && g.slot_of[type] >= 0, "type has no map");
	require(!charge || g.charged, "the maps have no charge maps (tables without charge-dependent terms)");
	const size_t npts = size_t(g.dim[0]) * g.dim[1] * g.dim[2];
	std::vector<double> pts(npts);
	hip_check(hipMemcpyAsync(pts.data(), (charge ? g.qpoints : g.points).as<double>() + size_t(g.slot_of[type]) * npts,
					  npts * sizeof(double), hipMemcpyDeviceToHost, c->stream),
			"download grid");
	hip_check(hipStreamSynchronize(c->stream), "download grid");
	// stored precision: round through it
	for (size_t k = 0; k < npts; k++)
		out[k] = g.precision == SGPU_FP64 ? pts[k] : double(float(pts[k]));
}
} // namespace

sgpu_status sgpu_download_charge_grid(sgpu_ctx* c, int grid_id, int type, double* out)
{
	return guarded(c, [&]() { download_map(c, grid_id, type, true, out); });
}

sgpu_status sgpu_download_grid(sgpu_ctx* c, int grid_id, int type, double* out)
{
	return guarded(c, [&]() { download_map(c, grid_id, type, false, out); });
}

sgpu_status sgpu_release_grid(sgpu_ctx* c, int grid_id)
{
	return guarded(c, [&]() {
		find_grid(c, grid_id);
		c->grids.erase(grid_id);
	});
}

sgpu_status sgpu_upload_models(sgpu_ctx* c, const sgpu_flat_model* models, int n, int precision, int* batch_id)
{
	return guarded(c, [&]() {
		require(models && n > 0 && batch_id, "no models");
		require(precision == SGPU_FP32 || precision == SGPU_FP64 || precision == SGPU_MIXED, "invalid precision");
		std::unique_ptr<Batch> b(new Batch);
		b->precision = precision;
		b->n = n;
		pack_batch(models, n, b->h);
		b->h.plan.bytes = plan_bytes(b->h.plan, precision, 1);
		require(plan_bytes(b->h.plan, SGPU_FP64, 1) <= LDS_LIMIT && b->h.plan.bytes <= LDS_LIMIT,
				"model too large for one workgroup's LDS");
		upload_batch(*b, c->stream);
		if (precision != SGPU_FP64)
		{
			b->fp64.reset(new Batch);
			b->fp64->precision = SGPU_FP64;
			b->fp64->n = n;
			b->fp64->h = b->h;
			b->fp64->h.plan.bytes = plan_bytes(b->h.plan, SGPU_FP64, 1);
			upload_batch(*b->fp64, c->stream);
		}
		*batch_id = c->next_id++;
		c->batches[*batch_id] = std::move(b);
	});
}

sgpu_status sgpu_release_models(sgpu_ctx* c, int batch_id)
{
	return guarded(c, [&]() {
		find_batch(c, batch_id);
		c->batches.erase(batch_id);
	});
}

} // extern "C"

namespace
{

template<typename T>
void run_eval(sgpu_ctx* c, const sgpu_eval_request* q)
{
	Batch& b = find_batch(c, q->batch_id);
	const int n = q->nconfs;
	require(n > 0 && q->model && q->conf && q->e, "eval request incomplete");
	require(q->conf_stride >= 1, "conf stride");
	for (int i = 0; i < n; i++)
	{
		require(q->model[i] >= 0 && q->model[i] < b.n, "model index out of range");
		require(q->conf_stride >= b.h.hdr[q->model[i]].conf_size, "conf stride below conf size");
		if (q->g)
			require(q->change_stride >= b.h.hdr[q->model[i]].change_size, "change stride below change size");
		if (q->coords)
			require(q->coord_stride >= 3 * b.h.hdr[q->model[i]].natoms, "coord stride below 3 * atoms");
	}
	EvalCtx<T> ec{};
	ec.tab = table_ptrs<T>(c);
	attach_user_grid(c, ec);
	for (int d = 0; d < 3; d++)
		ec.v[d] = T(q->v[d]);
	ensure_scratch(c->err_flag, sizeof(int));
	hip_check(hipMemsetAsync(c->err_flag.p, 0, sizeof(int), c->stream), "memset");
	ec.error = c->err_flag.as<int>();
	if (q->grid_id >= 0)
	{
		Grid& g = find_grid(c, q->grid_id);
		require(g.precision == eval_precision(b.precision), "grid and model precision differ");
		check_maps(b, g, q->model, n);
		ec.backend = 0;
		ec.grid = grid_ptrs<T>(g, q->slope);
		ec.slope = T(q->slope);
	}
	else
	{
		check_gd(q->gd);
		Receptor& r = find_receptor(c, q->receptor_id);
		Cells& cl = receptor_cells(c, r, *q->gd);
		ec.backend = 1;
		ec.rec = receptor_ptrs<T>(r, cl);
		ec.slope = T(q->slope);
	}
	const int cs = b.h.plan.cmax;
	std::vector<T> conf(size_t(n) * cs, T(0));
	for (int i = 0; i < n; i++)
		for (int k = 0; k < b.h.hdr[q->model[i]].conf_size; k++)
			conf[size_t(i) * cs + k] = T(q->conf[size_t(i) * q->conf_stride + k]);
	const int gs = b.h.plan.gmax, xs = 3 * b.h.plan.nmax, ps = b.h.plan.pmax;
	DBuf dm, dc, de, dg, dx;
	upload(dm, std::vector<int>(q->model, q->model + n), c->stream);
	upload(dc, conf, c->stream);
	de.alloc(size_t(n) * sizeof(T));
	dg.alloc(size_t(n) * gs * sizeof(T));
	dx.alloc(size_t(n) * xs * sizeof(T));
	ensure_scratch(c->scratch, size_t(n) * ps * sizeof(T));
	Timer tm(c, "eval");
	hip_check(launch_eval<T>(ptrs<T>(b), ec, b.h.plan, n, dm.as<int>(), dc.as<T>(), cs, de.as<T>(), dg.as<T>(), gs,
					  dx.as<T>(), xs, c->scratch.as<T>(), ps, c->stream),
			"eval kernel");
	tm.stop();
	std::vector<T> e(n), g(size_t(n) * gs), x(size_t(n) * xs);
	int flag = 0;
	hip_check(hipMemcpyAsync(e.data(), de.p, e.size() * sizeof(T), hipMemcpyDeviceToHost, c->stream), "download");
	hip_check(hipMemcpyAsync(g.data(), dg.p, g.size() * sizeof(T), hipMemcpyDeviceToHost, c->stream), "download");
	hip_check(hipMemcpyAsync(x.data(), dx.p, x.size() * sizeof(T), hipMemcpyDeviceToHost, c->stream), "download");
	hip_check(hipMemcpyAsync(&flag, c->err_flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream), "download");
	hip_check(hipStreamSynchronize(c->stream), "eval");
	for (int i = 0; i < n; i++)
	{
		const ModelHdr& h = b.h.hdr[q->model[i]];
		q->e[i] = double(e[i]);
		if (q->g)
			for (int k = 0; k < h.change_size; k++)
				q->g[size_t(i) * q->change_stride + k] = double(g[size_t(i) * gs + k]);
		if (q->coords)
			for (int k = 0; k < 3 * h.natoms; k++)
				q->coords[size_t(i) * q->coord_stride + k] = double(x[size_t(i) * xs + k]);
	}
	if (flag)
		throw Fail{SGPU_E_ATOM_OVERLAP, "Ligand atom exactly overlaps receptor atom.  I can't deal with this."};
}

// the size class of a model in a Monte Carlo launch, by its pair count
int mc_class(const ModelHdr& h)
{
	const int np = h.npairs;
	return np <= 256 ? 0 : np <= 512 ? 1 : np <= 1024 ? 2 : 3;
}

// the Monte Carlo launch's work-queue counters, cleared by a kernel (a
// memset node would sit in the captured graph)
__global__ void mc_queue_init(int* q, int n)
{
	const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
	if (i < n)
		q[i] = 0;
}

// stop_agreeing on the device, one thread per model: the best energy of
// each search unit (the lowest over its chains' bests), the lowest of them,
// and the units within eps of it; with at least need[m] agreeing, the model
// stops for good (its chains' step limits become 0)
template<typename S>
__global__ void early_stop_kernel(int nmodels, const int* mstart, const int* mchain, const int* munit,
		const int* nunits, const int* need, const char* chains, size_t state_bytes, size_t off_best, S eps,
		int* stopped, int* steps)
{
	const int m = int(blockIdx.x * blockDim.x + threadIdx.x);
	if (m >= nmodels || stopped[m] || nunits[m] < 2)
		return;
	auto best = [&](int k) { return *reinterpret_cast<const S*>(chains + size_t(mchain[k]) * state_bytes + off_best); };
	S lowest = 0;
	bool have = false;
	for (int u = 0; u < nunits[m]; u++)
	{
		S ub = 0;
		bool any = false;
		for (int k = mstart[m]; k < mstart[m + 1]; k++)
			if (munit[k] == u)
			{
				const S b = best(k);
				ub = any ? (b < ub ? b : ub) : b;
				any = true;
			}
		if (any)
		{
			lowest = have ? (ub < lowest ? ub : lowest) : ub;
			have = true;
		}
	}
	int agree = 0;
	for (int u = 0; u < nunits[m]; u++)
	{
		S ub = 0;
		bool any = false;
		for (int k = mstart[m]; k < mstart[m + 1]; k++)
			if (munit[k] == u)
			{
				const S b = best(k);
				ub = any ? (b < ub ? b : ub) : b;
				any = true;
			}
		if (any && ub - lowest <= eps)
			agree++;
	}
	if (agree >= need[m])
	{
		stopped[m] = 1;
		for (int k = mstart[m]; k < mstart[m + 1]; k++)
			steps[mchain[k]] = 0;
	}
}

// the occupancy rule of a size class: chains per workgroup for the most
// resident waves per CU (registers and LDS, from the runtime's occupancy
// calculator), then the larger workgroup; 0 when nothing fits
template<typename T, typename S>
int occupancy_waves(const WavePlan& plan, int precision, int variant, int maxper, int* resident)
{
	int waves = 0, best = 0;
	for (int w = 1; w <= std::min(MAX_CHAINS_PER_GROUP, maxper); w++)
	{
		const int bytes = plan_bytes(plan, precision, w);
		if (bytes + LDS_STATIC > LDS_LIMIT)
			break;
		const int r = w * mc_blocks_per_cu<T, S>(w, size_t(bytes), variant);
		if (r >= best && r > 0)
		{
			best = r;
			waves = w;
		}
	}
	if (resident)
		*resident = best;
	return waves;
}

// the default form: the persistent grid unless its kernel spills where the
// one-shot kernel does not; SGPU_QUEUE=0 / 1 forces either
template<typename T, typename S>
bool persist_default(int variant)
{
	const char* qenv = std::getenv("SGPU_QUEUE");
	return qenv ? std::atoi(qenv) != 0 : mc_private_bytes<T, S>(variant, true) <= mc_private_bytes<T, S>(variant, false);
}

template<typename T, typename S>
void run_mc(sgpu_ctx* c, int batch_id, int grid_id, const sgpu_mc_params* p, const sgpu_chains* ch,
		sgpu_chain_results* out)
{
	Batch& b = find_batch(c, batch_id);
	Grid& g = find_grid(c, grid_id);
	require(g.precision == eval_precision(b.precision), "grid and model precision differ");
	require(p && ch && out && ch->nchains > 0, "mc request incomplete");
	require(p->num_saved_mins >= 1 && p->num_saved_mins <= MAX_POSES, "num_saved_mins out of range");
	require(out->max_poses >= p->num_saved_mins, "result arrays hold fewer poses than num_saved_mins");
	const int n = ch->nchains;
	for (int i = 0; i < n; i++)
	{
		require(ch->model[i] >= 0 && ch->model[i] < b.n, "chain model out of range");
		const ModelHdr& h = b.h.hdr[ch->model[i]];
		require(ch->conf_stride >= h.conf_size && out->conf_stride >= h.conf_size, "conf stride");
		require(out->coord_stride >= 3 * h.nheavy, "coord stride");
		require(ch->num_steps[i] >= 0 && ch->maxiters[i] >= 0, "steps / maxiters");
		require(ch->rng[size_t(i) * 625 + 624] <= 624, "rng position");
		require(!ch->unit || ch->unit[i] >= 0, "search unit");
	}
	const bool early_stop = p->early_stop_checkpoints > 0;
	if (early_stop)
		require(p->early_stop_first >= 1 && p->early_stop_fraction > 0 && p->early_stop_fraction <= 1 &&
						p->early_stop_eps >= 0,
				"early stopping rule");
	check_maps(b, g, ch->model, n);
	// Launch plan.  Ligands fall into size classes by pair count; each class
	// gets its own LDS plan (sized for its largest member) and runs on its own
	// stream, concurrently with the others.  Within a class, workgroups hold
	// `waves` chains of one ligand (the largest power of two up to 8 that the
	// chains per ligand and 160 KB of LDS allow), costliest first.
	std::map<int, std::vector<int>> by_model;
	for (int i = 0; i < n; i++)
		by_model[ch->model[i]].push_back(i);
	auto class_of = [&](int mi) { return mc_class(b.h.hdr[mi]); };
	struct Blk
	{
		double cost;
		int model;
		std::vector<int> chains;
	};
	struct Cls
	{
		std::vector<int> models;
		WavePlan plan;
		int waves = 1;
		int resident = 0; // waves per CU
		int grid = 0;     // workgroups of the class's launch
		bool persist = false; // persistent grid over a work queue
		bool tuned = false;   // waves and form from the autotuner
		bool spec = false;    // one speculative chain per workgroup of `waves` waves
		std::vector<Blk> blks;
		size_t block_off = 0, slot_off = 0;
	};
	Cls cls[sgpu_ctx::NCLASS];
	for (auto& kv : by_model)
		cls[class_of(kv.first)].models.push_back(kv.first);
	hipDeviceProp_t prop;
	hip_check(hipGetDeviceProperties(&prop, c->device), "hipGetDeviceProperties");
	const int variant = (c->tab.kind == 1 || c->tab.ncomp == 4) ? 2 : p->visited ? 1 : 0;
	// Speculative chains: with fewer chains than the device has wave slots,
	// each chain gets `spec` waves (wave 0 runs it, the others evaluate the
	// next trials of its fast line search), without QuickVina 2's history;
	// the widest of 8, 4 and 2 whose n x width waves all fit at once.
	// SGPU_SPEC=1 turns it off, 2..8 forces the width.
	int spec = 1;
	if (!p->visited && !p->accurate_line_search)
	{
		const char* senv = std::getenv("SGPU_SPEC");
		const int forced = senv ? std::atoi(senv) : 0;
		for (int w : {8, 4, 3, 2})
		{
			if (forced ? w != forced : w == 3)
				continue;
			int slots = std::numeric_limits<int>::max();
			for (Cls& k : cls)
				if (!k.models.empty())
				{
					const WavePlan pl = plan_for(b.h, k.models);
					const int bytes = plan_bytes(pl, b.precision, w);
					const int per_cu = bytes + spec_static_bytes<S>() <= LDS_LIMIT
							? mc_blocks_per_cu<T, S>(w, size_t(bytes), variant, false, true)
							: 0;
					slots = std::min(slots, per_cu * w * prop.multiProcessorCount);
				}
			if (forced ? slots > 0 : int64_t(n) * w <= int64_t(slots))
			{
				spec = w;
				break;
			}
		}
	}
	ChainLayout L{};
	L.conf_stride = b.h.plan.cmax;
	int hmax = 1;
	for (const ModelHdr& h : b.h.hdr)
		hmax = std::max(hmax, h.nheavy);
	L.coord_stride = 3 * hmax;
	L.atom_stride = 3 * b.h.plan.nmax;
	L.pair_stride = b.h.plan.pmax;
	L.max_poses = MAX_POSES;
	size_t off = 0;
	auto take = [&](size_t bytes) {
		off = (off + 63) / 64 * 64;
		size_t o = off;
		off += bytes;
		return o;
	};
	L.off_mt = take(MT_N * sizeof(uint32_t));
	L.off_ints = take((3 + MAX_POSES) * sizeof(int));
	L.off_tmp = take(L.conf_stride * sizeof(S));
	L.off_scal = take(2 * sizeof(S));
	L.off_oute = take(MAX_POSES * sizeof(S));
	L.off_outc = take(size_t(MAX_POSES) * L.conf_stride * sizeof(S));
	L.off_outx = take(size_t(MAX_POSES) * L.coord_stride * sizeof(T));
	L.off_mc = take(L.atom_stride * sizeof(T));
	L.off_ro = take(MAX_TREES * 3 * sizeof(T));
	L.off_pbuf = take((sizeof(T) == 8 ? size_t(L.pair_stride) : 1) * sizeof(T));
	// the ordered pair loop's scratch of a speculative chain's helper waves
	L.off_hpbuf = take((sizeof(T) == 8 && spec > 1 ? size_t(spec - 1) * L.pair_stride : 1) * sizeof(T));
	L.off_ev = take(sizeof(int64_t));
	// QuickVina 2's history: 5n points of n coordinates, n up to the batch's
	// largest change size and at most 64 (chain_layout,
	// src/gpu/gpu_mc_structs.h:97-152)
	L.vis_n = p->visited ? std::min(b.h.plan.gmax, 64) : 0;
	L.off_vx = take(size_t(5) * L.vis_n * L.vis_n * sizeof(S));
	L.off_vf = take(size_t(5) * L.vis_n * sizeof(S));
	L.off_vb = take(size_t(10) * L.vis_n * sizeof(uint64_t));
	L.off_vs = take(3 * sizeof(int));
	L.bytes = (off + 63) / 64 * 64;
	// initial chain state (the history empty: zeros)
	std::vector<char> init(L.bytes * n, 0);
	for (int i = 0; i < n; i++)
	{
		char* base = init.data() + L.bytes * i;
		std::memcpy(base + L.off_mt, ch->rng + size_t(i) * 625, MT_N * sizeof(uint32_t));
		int* ints = reinterpret_cast<int*>(base + L.off_ints);
		ints[0] = int(ch->rng[size_t(i) * 625 + 624]);
		ints[1] = 0;
		ints[2] = 0;
		S* tmp = reinterpret_cast<S*>(base + L.off_tmp);
		for (int k = 0; k < b.h.hdr[ch->model[i]].conf_size; k++)
			tmp[k] = S(ch->conf[size_t(i) * ch->conf_stride + k]);
		S* scal = reinterpret_cast<S*>(base + L.off_scal);
		scal[0] = 0;
		scal[1] = Real_host_max<S>();
	}
	std::vector<int> block_model, slot_chain;
	for (Cls& k : cls)
	{
		if (k.models.empty())
			continue;
		k.plan = plan_for(b.h, k.models);
		int maxper = 1;
		for (int mi : k.models)
			maxper = std::max(maxper, int(by_model[mi].size()));
		// chains per workgroup by the occupancy rule
		k.waves = occupancy_waves<T, S>(k.plan, b.precision, variant, maxper, &k.resident);
		require(k.waves > 0, "chain state exceeds LDS");
		// a probe's override, else the tuned plan of this class, where it fits
		const int ki = int(&k - cls);
		const sgpu_ctx::ClassPlan* cp = nullptr;
		if (c->override_on[ki])
			cp = &c->plan_override[ki];
		else
		{
			auto it = c->tuned.find({b.precision, variant, ki});
			if (it != c->tuned.end())
				cp = &it->second;
		}
		if (spec > 1)
		{
			// one chain per workgroup of `spec` waves
			k.waves = spec;
			k.spec = true;
			k.resident = spec * mc_blocks_per_cu<T, S>(spec, size_t(plan_bytes(k.plan, b.precision, spec)), variant,
					false, true);
		}
		else if (cp && cp->waves >= 1 && cp->waves <= std::min(MAX_CHAINS_PER_GROUP, maxper) &&
				plan_bytes(k.plan, b.precision, cp->waves) + LDS_STATIC <= LDS_LIMIT &&
				mc_blocks_per_cu<T, S>(cp->waves, size_t(plan_bytes(k.plan, b.precision, cp->waves)), variant,
						cp->persist) > 0)
		{
			k.waves = cp->waves;
			k.persist = cp->persist;
			k.tuned = true;
		}
		if (std::getenv("SGPU_LOG_PLAN"))
		{
			const int bytes = plan_bytes(k.plan, b.precision, k.waves);
			std::fprintf(stderr,
					"[sgpu] mc class %d: %zu models, %d waves/workgroup%s, %d LDS bytes/workgroup, %d resident "
					"waves/CU (atoms %d, pairs %d, frames %d, change %d)\n",
					int(&k - cls), k.models.size(), k.waves, k.spec ? " (one speculative chain)" : "", bytes,
					k.resident, k.plan.nmax, k.plan.pmax, k.plan.fmax,
					k.plan.gmax);
		}
		for (int mi : k.models)
		{
			const std::vector<int>& cs = by_model[mi];
			for (size_t s = 0; s < cs.size(); s += size_t(k.spec ? 1 : k.waves))
			{
				Blk bk{0, mi, {}};
				for (size_t j = s; j < std::min(cs.size(), s + size_t(k.spec ? 1 : k.waves)); j++)
				{
					bk.chains.push_back(cs[j]);
					bk.cost = std::max(bk.cost, double(ch->num_steps[cs[j]]) *
									double(b.h.hdr[mi].npairs + 4 * b.h.hdr[mi].natoms));
				}
				k.blks.push_back(bk);
			}
		}
		std::stable_sort(k.blks.begin(), k.blks.end(), [](const Blk& x, const Blk& y) { return x.cost > y.cost; });
		k.block_off = block_model.size();
		k.slot_off = slot_chain.size();
		for (const Blk& bk : k.blks)
		{
			block_model.push_back(bk.model);
			for (int j = 0; j < k.waves; j++)
				slot_chain.push_back(k.spec ? bk.chains[0] : j < int(bk.chains.size()) ? bk.chains[j] : -1);
		}
	}
	DBuf dchains, dbm, dsc, dsteps, dmax;
	upload(dchains, init, c->stream);
	upload(dbm, block_model, c->stream);
	upload(dsc, slot_chain, c->stream);
	upload(dsteps, std::vector<int>(ch->num_steps, ch->num_steps + n), c->stream);
	upload(dmax, std::vector<int>(ch->maxiters, ch->maxiters + n), c->stream);
	EvalCtx<T> ec{};
	ec.tab = table_ptrs<T>(c);
	attach_user_grid(c, ec);
	ec.backend = 0;
	ec.grid = grid_ptrs<T>(g, p->cache_slope);
	ec.slope = T(p->cache_slope);
	McParams P{};
	P.temperature = p->temperature;
	P.amplitude = p->mutation_amplitude;
	P.min_rmsd = p->min_rmsd;
	P.num_saved = p->num_saved_mins;
	for (int d = 0; d < 3; d++)
	{
		P.hunt[d] = p->hunt_cap[d];
		P.authentic[d] = p->authentic_v[d];
	}
	P.accurate = p->accurate_line_search;
	P.early_term = p->early_term;
	P.visited = p->visited;
	int max_steps = 0;
	for (int i = 0; i < n; i++)
		max_steps = std::max(max_steps, ch->num_steps[i]);
	// Heaviest class first, on the context's stream; each further class on a
	// stream of its own (four streams at most: one hardware queue each, so
	// no launch queues behind another stream's work).  The waits that join
	// the streams are enqueued after every launch.  With progress reporting
	// the steps run as consecutive parts, joined and reported one by one;
	// with early stopping (run_monte_carlo, src/gpu/gpu_kernels_impl.cuh:
	// 742-808) the parts end at first/checkpoints, (first+1)/checkpoints, ..
	// of every chain's steps, and the rule runs after each part but the last.
	int parts = std::max(1, std::min(c->progress_chunks, max_steps)), part0 = 1;
	if (early_stop)
	{
		parts = p->early_stop_checkpoints;
		part0 = std::min(parts, p->early_stop_first);
	}
	// Persistent grids: each class's grid holds as many workgroups as the
	// device keeps resident at once for its plan; they take the class's
	// items (costliest first) from a work-queue counter, one counter per
	// class and part, cleared by a kernel at the start.
	// the default form unless a tuned plan says otherwise
	// SGPU_PERSIST_GRID caps a persistent grid's workgroups (each then runs
	// more items)
	const bool pdef = persist_default<T, S>(variant);
	const char* genv = std::getenv("SGPU_PERSIST_GRID");
	const int grid_cap = genv ? std::max(1, std::atoi(genv)) : std::numeric_limits<int>::max();
	for (Cls& k : cls)
		if (!k.blks.empty())
		{
			if (!k.tuned)
				k.persist = pdef && !k.spec;
			const int per_cu = mc_blocks_per_cu<T, S>(k.waves, size_t(plan_bytes(k.plan, b.precision, k.waves)),
					variant, k.persist, k.spec);
			k.grid = k.persist ? std::max(1, std::min({int(k.blks.size()), per_cu * prop.multiProcessorCount, grid_cap}))
			                   : int(k.blks.size());
		}
	DBuf dqueue;
	const int nqueue = sgpu_ctx::NCLASS * (parts + 1);
	dqueue.alloc(size_t(nqueue) * sizeof(int));
	// early stopping on the device (stop_agreeing, src/gpu/gpu_kernels_impl.cuh:
	// 607-657): per model, its chains grouped by search unit, the units' count
	// and the agreeing units needed, max(2, ceil(fraction x units))
	DBuf dms, dmc, dmu, dnu, dneed, dstop;
	if (early_stop)
	{
		std::vector<int> mstart(size_t(b.n) + 1, 0), mchain, munit, nunits(size_t(b.n), 0), need(size_t(b.n), 0);
		for (int mi = 0; mi < b.n; mi++)
		{
			mstart[size_t(mi)] = int(mchain.size());
			auto it = by_model.find(mi);
			if (it == by_model.end())
				continue;
			std::map<int, int> dense;
			int seq = 0;
			for (int ci : it->second)
			{
				const int u = ch->unit ? ch->unit[ci] : seq++;
				auto d = dense.find(u);
				if (d == dense.end())
					d = dense.emplace(u, int(dense.size())).first;
				mchain.push_back(ci);
				munit.push_back(d->second);
			}
			nunits[size_t(mi)] = int(dense.size());
			need[size_t(mi)] =
					std::max(2, int(std::ceil(p->early_stop_fraction * double(dense.size()) - 1e-9)));
		}
		mstart[size_t(b.n)] = int(mchain.size());
		upload(dms, mstart, c->stream);
		upload(dmc, mchain, c->stream);
		upload(dmu, munit, c->stream);
		upload(dnu, nunits, c->stream);
		upload(dneed, need, c->stream);
		upload(dstop, std::vector<int>(size_t(b.n), 0), c->stream);
	}
	// the progress of each part, for host nodes that report it
	struct Progress
	{
		sgpu_ctx* c;
		double frac;
	};
	std::vector<Progress> prog(size_t(parts) + 1);
	auto report = [](void* user) {
		const Progress* pr = static_cast<const Progress*>(user);
		pr->c->progress_fn(pr->c->progress_user, pr->frac);
	};
	// The whole search (every part, the early-stopping rule between parts,
	// the progress reports) is captured once into a HIP graph and launched
	// as one: no host round trip between parts.
	const bool use_graph = !std::getenv("SGPU_NO_GRAPH");
	Timer tm(c, "mc");
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	if (use_graph)
		hip_check(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed), "graph capture");
	auto enqueue = [&]() {
		hipLaunchKernelGGL(mc_queue_init, dim3(unsigned((nqueue + 255) / 256)), dim3(256), 0, c->stream,
				dqueue.as<int>(), nqueue);
		hip_check(hipGetLastError(), "queue init");
		for (int part = part0; part <= parts; part++)
		{
			const int lo = part == part0 ? 0 : part - 1, hi = part;
			// everything enqueued on the main stream so far (the chain upload,
			// and after a checkpoint the early-stop step limits) comes before
			// this part's launches on every stream
			hip_check(hipEventRecord(c->part_start, c->stream), "event");
			std::vector<hipStream_t> side;
			for (int k = sgpu_ctx::NCLASS - 1; k >= 0; k--)
			{
				const Cls& cl = cls[k];
				if (cl.blks.empty())
					continue;
				hipStream_t st = c->stream;
				bool first = true;
				for (int j = k + 1; j < sgpu_ctx::NCLASS; j++)
					if (!cls[j].blks.empty())
						first = false;
				if (!first)
				{
					st = c->class_stream[side.size()];
					side.push_back(st);
					hip_check(hipStreamWaitEvent(st, c->part_start, 0), "stream wait");
				}
				int* q = cl.persist ? dqueue.as<int>() + size_t(part) * sgpu_ctx::NCLASS + size_t(k) : nullptr;
				hip_check(launch_mc<T, S>(ptrs<T>(b), ec, cl.plan, P, L, dchains.as<char>(), int(cl.blks.size()),
								  cl.waves, dbm.as<int>() + cl.block_off, dsc.as<int>() + cl.slot_off,
								  dsteps.as<int>(), dmax.as<int>(), lo, hi, parts, cl.grid, q, cl.spec, st),
						"mc kernel");
			}
			for (size_t j = 0; j < side.size(); j++)
			{
				hip_check(hipEventRecord(c->class_done[j], side[j]), "event");
				hip_check(hipStreamWaitEvent(c->stream, c->class_done[j], 0), "stream wait");
			}
			if (part < parts && early_stop)
			{
				hipLaunchKernelGGL((early_stop_kernel<S>), dim3(unsigned((b.n + 63) / 64)), dim3(64), 0, c->stream,
						b.n, dms.as<int>(), dmc.as<int>(), dmu.as<int>(), dnu.as<int>(), dneed.as<int>(),
						dchains.as<char>(), L.bytes, L.off_scal + sizeof(S), S(p->early_stop_eps), dstop.as<int>(),
						dsteps.as<int>());
				hip_check(hipGetLastError(), "early stop kernel");
			}
			if (part < parts && c->progress_fn)
			{
				prog[size_t(part)] = {c, double(hi) / double(parts)};
				hip_check(hipLaunchHostFunc(c->stream, report, &prog[size_t(part)]), "progress");
			}
		}
	};
	try
	{
		enqueue();
	}
	catch (...)
	{
		if (use_graph)
		{
			(void)hipStreamEndCapture(c->stream, &graph);
			if (graph)
				(void)hipGraphDestroy(graph);
		}
		throw;
	}
	if (use_graph)
	{
		hip_check(hipStreamEndCapture(c->stream, &graph), "graph capture");
		const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
		if (ie == hipSuccess)
		{
			const hipError_t le = hipGraphLaunch(exec, c->stream);
			(void)hipGraphExecDestroy(exec);
			(void)hipGraphDestroy(graph);
			hip_check(le, "graph launch");
		}
		else
		{
			(void)hipGraphDestroy(graph);
			hip_check(ie, "graph instantiate");
		}
	}
	tm.stop();
	if (c->progress_fn)
		c->progress_fn(c->progress_user, 1.0);
	hip_check(hipMemcpyAsync(init.data(), dchains.p, init.size(), hipMemcpyDeviceToHost, c->stream), "download");
	hip_check(hipStreamSynchronize(c->stream), "mc");
	for (int i = 0; i < n; i++)
	{
		const char* base = init.data() + L.bytes * i;
		const ModelHdr& h = b.h.hdr[ch->model[i]];
		const int* ints = reinterpret_cast<const int*>(base + L.off_ints);
		const int cnt = std::min(ints[1], out->max_poses);
		const S* oe = reinterpret_cast<const S*>(base + L.off_oute);
		const S* oc = reinterpret_cast<const S*>(base + L.off_outc);
		const T* ox = reinterpret_cast<const T*>(base + L.off_outx);
		out->count[i] = cnt;
		for (int k = 0; k < cnt; k++)
		{
			const int slot = ints[3 + k];
			out->e[size_t(i) * out->max_poses + k] = double(oe[slot]);
			for (int j = 0; j < h.conf_size; j++)
				out->conf[(size_t(i) * out->max_poses + k) * out->conf_stride + j] = double(oc[size_t(slot) * L.conf_stride + j]);
			for (int j = 0; j < 3 * h.nheavy; j++)
				out->coords[(size_t(i) * out->max_poses + k) * out->coord_stride + j] =
						double(ox[size_t(slot) * L.coord_stride + j]);
		}
		if (out->evals)
			out->evals[i] = *reinterpret_cast<const int64_t*>(base + L.off_ev);
	}
}

// The launch-plan autotuner: for each size class of the request, the
// class's chains alone run probe_steps Monte Carlo steps from their start
// state under every candidate plan (chains per workgroup 1, 2, 4, 8 where
// LDS and the chains per ligand allow; one-shot or persistent grid), timed
// with HIP events; the fastest plan is kept for the context's later
// launches of that precision, kernel variant and class.  The chains'
// trajectories do not depend on the plan, so the probes time the same work.
template<typename T, typename S>
void run_tune(sgpu_ctx* c, int batch_id, int grid_id, const sgpu_mc_params* p, const sgpu_chains* ch,
		int probe_steps, sgpu_mc_tune_report* rep)
{
	Batch& b = find_batch(c, batch_id);
	require(p && ch && ch->nchains > 0 && probe_steps >= 1 && rep, "tune request incomplete");
	const int variant = (c->tab.kind == 1 || c->tab.ncomp == 4) ? 2 : p->visited ? 1 : 0;
	std::vector<int> members[sgpu_ctx::NCLASS];
	int hmax = 1;
	for (int i = 0; i < ch->nchains; i++)
	{
		require(ch->model[i] >= 0 && ch->model[i] < b.n, "chain model out of range");
		members[mc_class(b.h.hdr[ch->model[i]])].push_back(i);
		hmax = std::max(hmax, b.h.hdr[ch->model[i]].nheavy);
	}
	*rep = sgpu_mc_tune_report{};
	sgpu_mc_params probe = *p;
	probe.early_stop_checkpoints = 0;
	const sgpu_progress_fn saved_fn = c->progress_fn;
	c->progress_fn = nullptr;
	struct Restore
	{
		sgpu_ctx* c;
		sgpu_progress_fn fn;
		~Restore()
		{
			c->progress_fn = fn;
			for (bool& o : c->override_on)
				o = false;
		}
	} restore{c, saved_fn};
	// per class: the largest chains per workgroup up to a target that its
	// LDS plan, its chains per ligand and the kernel form allow
	struct ClassInfo
	{
		WavePlan plan;
		int maxper = 1;
	};
	ClassInfo info[sgpu_ctx::NCLASS];
	for (int k = 0; k < sgpu_ctx::NCLASS; k++)
	{
		if (members[k].empty())
			continue;
		std::map<int, int> per_model;
		std::vector<int> models;
		for (int i : members[k])
			if (per_model[ch->model[i]]++ == 0)
				models.push_back(ch->model[i]);
		info[k].plan = plan_for(b.h, models);
		for (auto& kv : per_model)
			info[k].maxper = std::max(info[k].maxper, kv.second);
		rep->class_index[rep->nclasses++] = k;
	}
	auto fit = [&](int k, int target, bool pers) {
		for (int w = std::min({target, MAX_CHAINS_PER_GROUP, info[k].maxper}); w >= 1; w--)
		{
			const int bytes = plan_bytes(info[k].plan, b.precision, w);
			if (bytes + LDS_STATIC <= LDS_LIMIT && mc_blocks_per_cu<T, S>(w, size_t(bytes), variant, pers) > 0)
				return w;
		}
		return 0;
	};
	// joint candidates: one plan for every class at once, probed on the
	// whole batch, since the classes share the device (workgroups of one
	// class that leave part of a CU free can keep the larger workgroups of
	// another waiting).  The first is the default (the occupancy rule and
	// the default form), then a target of 8, 4, 2 or 1 chains per workgroup
	// in either form.
	struct Joint
	{
		sgpu_ctx::ClassPlan plan[sgpu_ctx::NCLASS];
	};
	std::vector<Joint> cands;
	{
		Joint d{};
		for (int k = 0; k < sgpu_ctx::NCLASS; k++)
			if (!members[k].empty())
				d.plan[k] = {occupancy_waves<T, S>(info[k].plan, b.precision, variant, info[k].maxper, nullptr),
						persist_default<T, S>(variant)};
		cands.push_back(d);
	}
	for (int target : {8, 4, 2, 1})
		for (int pers = 0; pers < 2; pers++)
		{
			Joint j{};
			bool ok = true;
			for (int k = 0; k < sgpu_ctx::NCLASS; k++)
				if (!members[k].empty())
				{
					j.plan[k] = {fit(k, target, pers != 0), pers != 0};
					ok = ok && j.plan[k].waves > 0;
				}
			bool dup = false;
			for (const Joint& o : cands)
			{
				bool same = true;
				for (int k = 0; k < sgpu_ctx::NCLASS; k++)
					same = same && o.plan[k].waves == j.plan[k].waves && o.plan[k].persist == j.plan[k].persist;
				dup = dup || same;
			}
			if (ok && !dup && int(cands.size()) < SGPU_TUNE_MAX_CANDIDATES)
				cands.push_back(j);
		}
	require(!cands.empty(), "no launch plan fits these chains");
	// each chain's steps scaled by one factor, the longest to probe_steps:
	// the probe keeps the batch's spread of chain lengths, which sets how
	// the classes' workgroups finish and refill the device
	const int n = ch->nchains;
	int longest = 1;
	for (int i = 0; i < n; i++)
		longest = std::max(longest, ch->num_steps[i]);
	std::vector<int> steps(n);
	for (int i = 0; i < n; i++)
		steps[size_t(i)] = std::max(1, int(std::ceil(double(ch->num_steps[i]) * std::min(probe_steps, longest) / longest)));
	sgpu_chains all = *ch;
	all.num_steps = steps.data();
	const int maxp = p->num_saved_mins;
	std::vector<int> cnt(n);
	std::vector<double> oe(size_t(n) * maxp), oc(size_t(n) * maxp * ch->conf_stride), ox(size_t(n) * maxp * 3 * hmax);
	sgpu_chain_results res{maxp, ch->conf_stride, 3 * hmax, cnt.data(), oe.data(), oc.data(), ox.data(), nullptr};
	double best_ms = 0;
	int best = 0;
	for (int j = -1; j < int(cands.size()); j++)
	{
		// j = -1: an untimed first run (code objects, buffers)
		const Joint& cj = cands[size_t(std::max(0, j))];
		for (int k = 0; k < sgpu_ctx::NCLASS; k++)
		{
			c->override_on[k] = !members[k].empty();
			c->plan_override[k] = cj.plan[k];
		}
		run_mc<T, S>(c, batch_id, grid_id, &probe, &all, &res);
		if (j < 0)
			continue;
		const double ms = c->last_ms["mc"];
		for (int ci = 0; ci < rep->nclasses; ci++)
		{
			const int k = rep->class_index[ci];
			rep->cand_waves[ci][j] = cj.plan[k].waves;
			rep->cand_persist[ci][j] = cj.plan[k].persist ? 1 : 0;
			rep->cand_ms[ci][j] = ms;
			rep->ncand[ci] = j + 1;
		}
		if (j == 0)
			best_ms = ms;
		// the default stays unless beaten by more than the probe's noise
		else if (ms < best_ms && ms < rep->cand_ms[0][0] * (1 - TUNE_MARGIN))
		{
			best_ms = ms;
			best = j;
		}
	}
	for (int ci = 0; ci < rep->nclasses; ci++)
	{
		const int k = rep->class_index[ci];
		rep->chosen_waves[ci] = cands[size_t(best)].plan[k].waves;
		rep->chosen_persist[ci] = cands[size_t(best)].plan[k].persist ? 1 : 0;
		c->tuned[{b.precision, variant, k}] = cands[size_t(best)].plan[k];
	}
}

template<typename T, typename S>
void run_refine(sgpu_ctx* c, const sgpu_refine_request* q)
{
	Batch& b = find_batch(c, q->batch_id);
	Receptor& r = find_receptor(c, q->receptor_id);
	check_gd(q->gd);
	const int n = q->nposes;
	require(n > 0 && q->model && q->conf && q->maxiters && q->e, "refine request incomplete");
	for (int i = 0; i < n; i++)
	{
		require(q->model[i] >= 0 && q->model[i] < b.n, "pose model out of range");
		require(q->conf_stride >= b.h.hdr[q->model[i]].conf_size, "conf stride");
		require(q->maxiters[i] >= 0, "maxiters");
		if (q->coords)
			require(q->coord_stride >= 3 * b.h.hdr[q->model[i]].nheavy, "coord stride");
	}
	DBuf dboxes;
	if (q->boxes)
	{
		for (int i = 0; i < n; i++)
			for (int d = 0; d < 3; d++)
			{
				const double lo = q->boxes[6 * size_t(i) + d], hi = q->boxes[6 * size_t(i) + 3 + d];
				require(lo < hi, "pose box end must exceed begin");
				require(lo >= q->gd->begin[d] && hi <= q->gd->end[d], "pose box outside the cell box");
			}
		upload(dboxes, std::vector<double>(q->boxes, q->boxes + 6 * size_t(n)), c->stream);
	}
	Cells& cl = receptor_cells(c, r, *q->gd);
	EvalCtx<T> ec{};
	ec.tab = table_ptrs<T>(c);
	attach_user_grid(c, ec);
	ec.backend = 1;
	ec.rec = receptor_ptrs<T>(r, cl);
	for (int d = 0; d < 3; d++)
		ec.v[d] = T(q->v[d]);
	ensure_scratch(c->err_flag, sizeof(int));
	hip_check(hipMemsetAsync(c->err_flag.p, 0, sizeof(int), c->stream), "memset");
	ec.error = c->err_flag.as<int>();
	const int cs = b.h.plan.cmax, ps = b.h.plan.pmax;
	int hmax = 1;
	for (const ModelHdr& h : b.h.hdr)
		hmax = std::max(hmax, h.nheavy);
	const int xs = 3 * hmax;
	std::vector<S> conf(size_t(n) * cs, S(0));
	for (int i = 0; i < n; i++)
		for (int k = 0; k < b.h.hdr[q->model[i]].conf_size; k++)
			conf[size_t(i) * cs + k] = S(q->conf[size_t(i) * q->conf_stride + k]);
	DBuf dm, dc, dmi, de, dx, dev, dtr;
	upload(dm, std::vector<int>(q->model, q->model + n), c->stream);
	upload(dc, conf, c->stream);
	upload(dmi, std::vector<int>(q->maxiters, q->maxiters + n), c->stream);
	de.alloc(size_t(n) * sizeof(S));
	dx.alloc(size_t(n) * xs * sizeof(T));
	dev.alloc(size_t(n) * sizeof(int64_t));
	dtr.alloc(size_t(n) * sizeof(int));
	ensure_scratch(c->scratch, size_t(n) * ps * sizeof(T));
	Timer tm(c, "refine");
	hip_check(launch_refine<T, S>(ptrs<T>(b), ec, b.h.plan, n, dm.as<int>(), dc.as<S>(), cs, dmi.as<int>(),
					  q->accurate_line_search, q->early_term, de.as<S>(), dx.as<T>(), xs, c->scratch.as<T>(), ps,
					  dev.as<int64_t>(), q->boxes ? dboxes.as<double>() : nullptr, dtr.as<int>(), c->stream),
			"refine kernel");
	tm.stop();
	std::vector<S> e(n);
	std::vector<T> x(size_t(n) * xs);
	int flag = 0;
	hip_check(hipMemcpyAsync(conf.data(), dc.p, conf.size() * sizeof(S), hipMemcpyDeviceToHost, c->stream), "download");
	hip_check(hipMemcpyAsync(e.data(), de.p, e.size() * sizeof(S), hipMemcpyDeviceToHost, c->stream), "download");
	hip_check(hipMemcpyAsync(x.data(), dx.p, x.size() * sizeof(T), hipMemcpyDeviceToHost, c->stream), "download");
	hip_check(hipMemcpyAsync(&flag, c->err_flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream), "download");
	if (q->tries)
		hip_check(hipMemcpyAsync(q->tries, dtr.p, size_t(n) * sizeof(int), hipMemcpyDeviceToHost, c->stream),
				"download");
	hip_check(hipStreamSynchronize(c->stream), "refine");
	for (int i = 0; i < n; i++)
	{
		const ModelHdr& h = b.h.hdr[q->model[i]];
		for (int k = 0; k < h.conf_size; k++)
			q->conf[size_t(i) * q->conf_stride + k] = double(conf[size_t(i) * cs + k]);
		q->e[i] = e[i] >= Real_host_max<S>() ? DBL_MAX : double(e[i]);
		if (q->coords)
			for (int k = 0; k < 3 * h.nheavy; k++)
				q->coords[size_t(i) * q->coord_stride + k] = double(x[size_t(i) * xs + k]);
	}
	if (flag)
		throw Fail{SGPU_E_ATOM_OVERLAP, "Ligand atom exactly overlaps receptor atom.  I can't deal with this."};
}

template<typename T, typename S>
void run_bfgs(sgpu_ctx* c, const sgpu_bfgs_request* q)
{
	Batch& b = find_batch(c, q->batch_id);
	const int n = q->nconfs;
	require(n > 0 && q->model && q->conf && q->maxiters && q->e, "bfgs request incomplete");
	for (int i = 0; i < n; i++)
	{
		require(q->model[i] >= 0 && q->model[i] < b.n, "conf model out of range");
		require(q->conf_stride >= b.h.hdr[q->model[i]].conf_size, "conf stride");
		require(q->maxiters[i] >= 0, "maxiters");
	}
	EvalCtx<T> ec{};
	ec.tab = table_ptrs<T>(c);
	attach_user_grid(c, ec);
	for (int d = 0; d < 3; d++)
		ec.v[d] = T(q->v[d]);
	ensure_scratch(c->err_flag, sizeof(int));
	hip_check(hipMemsetAsync(c->err_flag.p, 0, sizeof(int), c->stream), "memset");
	ec.error = c->err_flag.as<int>();
	ec.slope = T(q->slope);
	if (q->grid_id >= 0)
	{
		Grid& g = find_grid(c, q->grid_id);
		require(g.precision == eval_precision(b.precision), "grid and model precision differ");
		check_maps(b, g, q->model, n);
		ec.backend = 0;
		ec.grid = grid_ptrs<T>(g, q->slope);
	}
	else
	{
		check_gd(q->gd);
		Receptor& r = find_receptor(c, q->receptor_id);
		Cells& cl = receptor_cells(c, r, *q->gd);
		ec.backend = 1;
		ec.rec = receptor_ptrs<T>(r, cl);
	}
	const int cs = b.h.plan.cmax, ps = b.h.plan.pmax;
	std::vector<S> conf(size_t(n) * cs, S(0));
	for (int i = 0; i < n; i++)
		for (int k = 0; k < b.h.hdr[q->model[i]].conf_size; k++)
			conf[size_t(i) * cs + k] = S(q->conf[size_t(i) * q->conf_stride + k]);
	DBuf dm, dc, dmi, de, dev;
	upload(dm, std::vector<int>(q->model, q->model + n), c->stream);
	upload(dc, conf, c->stream);
	upload(dmi, std::vector<int>(q->maxiters, q->maxiters + n), c->stream);
	de.alloc(size_t(n) * sizeof(S));
	dev.alloc(size_t(n) * sizeof(int64_t));
	ensure_scratch(c->scratch, size_t(n) * ps * sizeof(T));
	Timer tm(c, "bfgs");
	hip_check(launch_bfgs<T, S>(ptrs<T>(b), ec, b.h.plan, n, dm.as<int>(), dc.as<S>(), cs, dmi.as<int>(),
					  q->accurate_line_search, q->early_term, de.as<S>(), c->scratch.as<T>(), ps, dev.as<int64_t>(),
					  c->stream),
			"bfgs kernel");
	tm.stop();
	std::vector<S> e(n);
	std::vector<int64_t> ev(n);
	int flag = 0;
	hip_check(hipMemcpyAsync(conf.data(), dc.p, conf.size() * sizeof(S), hipMemcpyDeviceToHost, c->stream), "download");
	hip_check(hipMemcpyAsync(e.data(), de.p, e.size() * sizeof(S), hipMemcpyDeviceToHost, c->stream), "download");
	hip_check(hipMemcpyAsync(ev.data(), dev.p, ev.size() * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream),
			"download");
	hip_check(hipMemcpyAsync(&flag, c->err_flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream), "download");
	hip_check(hipStreamSynchronize(c->stream), "bfgs");
	for (int i = 0; i < n; i++)
	{
		const ModelHdr& h = b.h.hdr[q->model[i]];
		for (int k = 0; k < h.conf_size; k++)
			q->conf[size_t(i) * q->conf_stride + k] = double(conf[size_t(i) * cs + k]);
		q->e[i] = double(e[i]);
		if (q->evals)
			q->evals[i] = ev[i];
	}
	if (flag)
		throw Fail{SGPU_E_ATOM_OVERLAP, "Ligand atom exactly overlaps receptor atom.  I can't deal with this."};
}

void run_score(sgpu_ctx* c, const sgpu_score_request* q)
{
	Batch& b0 = find_batch(c, q->batch_id);
	Batch& b = b0.precision == SGPU_FP64 ? b0 : *b0.fp64;
	Receptor& r = find_receptor(c, q->receptor_id);
	check_gd(q->gd);
	const int n = q->nconfs;
	require(n > 0 && q->model && q->conf && q->e && q->mode >= 0 && q->mode <= 2, "score request incomplete");
	for (int i = 0; i < n; i++)
	{
		require(q->model[i] >= 0 && q->model[i] < b.n, "model index out of range");
		require(q->conf_stride >= b.h.hdr[q->model[i]].conf_size, "conf stride");
	}
	Cells& cl = receptor_cells(c, r, *q->gd);
	EvalCtx<double> ec{};
	ec.tab = table_ptrs<double>(c);
	attach_user_grid(c, ec);
	ec.backend = 1;
	ec.rec = receptor_ptrs<double>(r, cl);
	ec.slope = q->slope;
	for (int d = 0; d < 3; d++)
		ec.v[d] = q->v[d];
	const int cs = b.h.plan.cmax, ps = std::max(b.h.plan.pmax, 64);
	std::vector<double> conf(size_t(n) * cs, 0.0);
	for (int i = 0; i < n; i++)
		for (int k = 0; k < b.h.hdr[q->model[i]].conf_size; k++)
			conf[size_t(i) * cs + k] = q->conf[size_t(i) * q->conf_stride + k];
	DBuf dm, dc, de;
	upload(dm, std::vector<int>(q->model, q->model + n), c->stream);
	upload(dc, conf, c->stream);
	de.alloc(size_t(n) * sizeof(double));
	ensure_scratch(c->scratch2, size_t(n) * ps * sizeof(double));
	int flag = 0;
	Timer tm(c, "score");
	if (q->mode == 2)
	{
		// model::eval_deriv's energy on the non_cache back end with the
		// search's tables: the function the refinement minimises, in FP64
		ensure_scratch(c->err_flag, sizeof(int));
		hip_check(hipMemsetAsync(c->err_flag.p, 0, sizeof(int), c->stream), "memset");
		ec.error = c->err_flag.as<int>();
		hip_check(launch_eval<double>(ptrs<double>(b), ec, b.h.plan, n, dm.as<int>(), dc.as<double>(), cs,
						  de.as<double>(), nullptr, 0, nullptr, 0, c->scratch2.as<double>(), ps, c->stream),
				"eval kernel");
		hip_check(hipMemcpyAsync(&flag, c->err_flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream), "download");
	}
	else
		hip_check(launch_score<double>(ptrs<double>(b), ec, b.h.plan, q->mode, n, dm.as<int>(), dc.as<double>(), cs,
						  de.as<double>(), c->scratch2.as<double>(), ps, c->stream),
				"score kernel");
	tm.stop();
	hip_check(hipMemcpyAsync(q->e, de.p, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, c->stream), "download");
	hip_check(hipStreamSynchronize(c->stream), "score");
	if (flag)
		throw Fail{SGPU_E_ATOM_OVERLAP, "Ligand atom exactly overlaps receptor atom.  I can't deal with this."};
}


template<typename T>
void run_eval_noalloc(sgpu_ctx* c, const sgpu_eval_request_dev* q)
{
	Batch& b = find_batch(c, q->batch_id);
	const int n = q->nconfs;
	require(n > 0 && q->model && q->conf && q->e, "eval request incomplete");
	const WavePlan& plan = b.h.plan;
	// the caller's strides cover the batch's largest model (its own sizes,
	// not the padded widths of the device plan)
	int csize = 0, gsize = 0, natoms = 0;
	for (const ModelHdr& h : b.h.hdr)
	{
		csize = std::max(csize, h.conf_size);
		gsize = std::max(gsize, h.change_size);
		natoms = std::max(natoms, h.natoms);
	}
	require(q->conf_stride >= csize, "conf stride below the batch's largest conf");
	require(!q->g || q->change_stride >= gsize, "change stride below the batch's largest change size");
	require(!q->coords || q->coord_stride >= 3 * natoms, "coord stride below 3 x the batch's largest model");
	const NoallocLayout L = noalloc_layout(n, plan.cmax, plan.gmax, plan.nmax, plan.pmax, sizeof(T));
	if (L.bytes > size_t(SGPU_NOALLOC_SCRATCH_BYTES))
		throw Fail{SGPU_E_INVALID, "request needs " + std::to_string(L.bytes) + " bytes of scratch, above the " +
						std::to_string(SGPU_NOALLOC_SCRATCH_BYTES) + " of the allocation-free entry point"};
	require(c->noalloc_scratch != nullptr, "hipGetSymbolAddress gave no scratch address for this device");
	EvalCtx<T> ec{};
	ec.tab = table_ptrs<T>(c);
	attach_user_grid(c, ec);
	for (int d = 0; d < 3; d++)
		ec.v[d] = T(q->v[d]);
	if (q->grid_id >= 0)
	{
		Grid& g = find_grid(c, q->grid_id);
		require(g.precision == eval_precision(b.precision), "grid and model precision differ");
		for (int mi = 0; mi < b.n; mi++)
			check_maps(b, g, &mi, 1);
		ec.backend = 0;
		ec.grid = grid_ptrs<T>(g, q->slope);
		ec.slope = T(q->slope);
	}
	else
	{
		require(q->gd != nullptr, "no box for the direct receptor term");
		check_gd(q->gd);
		Receptor& r = find_receptor(c, q->receptor_id);
		std::vector<double> key;
		for (int i = 0; i < 3; i++)
		{
			key.push_back(q->gd->begin[i]);
			key.push_back(q->gd->end[i]);
			key.push_back(q->gd->n[i]);
		}
		auto it = r.cells.find(key);
		require(it != r.cells.end(), "the receptor's cells for this box are not built: an sgpu_eval_confs, "
		                             "sgpu_run_refine or sgpu_score call with this box builds them");
		ec.backend = 1;
		ec.rec = receptor_ptrs<T>(r, *it->second);
		ec.slope = T(q->slope);
	}
	NoallocRequest nq{};
	nq.n = n;
	nq.model = q->model;
	nq.conf = q->conf;
	nq.conf_stride = q->conf_stride;
	nq.e = q->e;
	nq.g = q->g;
	nq.change_stride = q->change_stride;
	nq.coords = q->coords;
	nq.coord_stride = q->coord_stride;
	nq.status = q->status;
	hipStream_t st = q->stream ? static_cast<hipStream_t>(q->stream) : c->stream;
	hip_check(launch_eval_noalloc<T>(ptrs<T>(b), ec, plan, b.n, nq, c->noalloc_scratch,
					  size_t(SGPU_NOALLOC_SCRATCH_BYTES), st),
			"eval kernel");
	if (q->status)
		return;
	// no status word given: wait, and report the status as the call's
	int word = 0;
	hip_check(hipMemcpyAsync(&word, c->noalloc_scratch + L.status, sizeof(int), hipMemcpyDeviceToHost, st),
			"status");
	hip_check(hipStreamSynchronize(st), "eval");
	if (word & NOALLOC_BAD_MODEL)
		throw Fail{SGPU_E_INVALID, "model index out of range"};
	if (word & NOALLOC_ATOM_OVERLAP)
		throw Fail{SGPU_E_ATOM_OVERLAP, "Ligand atom exactly overlaps receptor atom.  I can't deal with this."};
}
} // namespace

extern "C" {

sgpu_status sgpu_eval_confs(sgpu_ctx* c, const sgpu_eval_request* q)
{
	return guarded(c, [&]() {
		require(q != nullptr, "no request");
		require(c->tab.ok, "tables not uploaded");
		Batch& b = find_batch(c, q->batch_id);
		if (b.precision == SGPU_FP64)
			run_eval<double>(c, q);
		else
			run_eval<float>(c, q);
	});
}

sgpu_status sgpu_run_bfgs(sgpu_ctx* c, const sgpu_bfgs_request* q)
{
	return guarded(c, [&]() {
		require(q != nullptr, "no request");
		require(c->tab.ok, "tables not uploaded");
		Batch& b = find_batch(c, q->batch_id);
		if (b.precision == SGPU_FP64)
			run_bfgs<double, double>(c, q);
		else if (b.precision == SGPU_MIXED)
			run_bfgs<float, double>(c, q);
		else
			run_bfgs<float, float>(c, q);
	});
}

sgpu_status sgpu_eval_confs_noalloc(sgpu_ctx* c, const sgpu_eval_request_dev* q)
{
	return guarded(c, [&]() {
		require(q != nullptr, "no request");
		require(c->tab.ok, "tables not uploaded");
		Batch& b = find_batch(c, q->batch_id);
		if (b.precision == SGPU_FP64)
			run_eval_noalloc<double>(c, q);
		else
			run_eval_noalloc<float>(c, q);
	});
}

sgpu_status sgpu_run_mc(sgpu_ctx* c, int batch_id, int grid_id, const sgpu_mc_params* p, const sgpu_chains* ch,
		sgpu_chain_results* out)
{
	return guarded(c, [&]() {
		require(c->tab.ok, "tables not uploaded");
		Batch& b = find_batch(c, batch_id);
		if (b.precision == SGPU_FP64)
			run_mc<double, double>(c, batch_id, grid_id, p, ch, out);
		else if (b.precision == SGPU_MIXED)
			run_mc<float, double>(c, batch_id, grid_id, p, ch, out);
		else
			run_mc<float, float>(c, batch_id, grid_id, p, ch, out);
	});
}

sgpu_status sgpu_tune_mc(sgpu_ctx* c, int batch_id, int grid_id, const sgpu_mc_params* p, const sgpu_chains* ch,
		int probe_steps, sgpu_mc_tune_report* out)
{
	return guarded(c, [&]() {
		require(c->tab.ok, "tables not uploaded");
		Batch& b = find_batch(c, batch_id);
		if (b.precision == SGPU_FP64)
			run_tune<double, double>(c, batch_id, grid_id, p, ch, probe_steps, out);
		else if (b.precision == SGPU_MIXED)
			run_tune<float, double>(c, batch_id, grid_id, p, ch, probe_steps, out);
		else
			run_tune<float, float>(c, batch_id, grid_id, p, ch, probe_steps, out);
	});
}

sgpu_status sgpu_run_refine(sgpu_ctx* c, const sgpu_refine_request* q)
{
	return guarded(c, [&]() {
		require(q != nullptr, "no request");
		require(c->tab.ok, "tables not uploaded");
		Batch& b = find_batch(c, q->batch_id);
		if (b.precision == SGPU_FP64)
			run_refine<double, double>(c, q);
		else if (b.precision == SGPU_MIXED)
			run_refine<float, double>(c, q);
		else
			run_refine<float, float>(c, q);
	});
}

sgpu_status sgpu_score(sgpu_ctx* c, const sgpu_score_request* q)
{
	return guarded(c, [&]() {
		require(q != nullptr, "no request");
		require(c->tab.ok, "tables not uploaded");
		run_score(c, q);
	});
}

sgpu_status sgpu_set_progress(sgpu_ctx* c, int chunks, sgpu_progress_fn fn, void* user)
{
	return guarded(c, [&]() {
		require(chunks >= 1 && chunks <= 1000, "1..1000 progress chunks");
		c->progress_chunks = chunks;
		c->progress_fn = fn;
		c->progress_user = user;
	});
}

sgpu_status sgpu_last_kernel_ms(sgpu_ctx* c, const char* kind, double* ms)
{
	return guarded(c, [&]() {
		require(kind && ms, "arguments");
		auto it = c->last_ms.find(kind);
		require(it != c->last_ms.end(), "no launch of that kind yet");
		*ms = it->second;
	});
}

} // extern "C"
