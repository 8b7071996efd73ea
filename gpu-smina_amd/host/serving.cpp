#include "serving.h"

#include <algorithm>
#include <chrono>
#include <stdexcept>

namespace sgpu_host
{

DockingServer::DockingServer(const ServerSettings& s, const sgpu_scoring& sf) : s_(s)
{
	if (s_.workers < 1 || s_.max_batch < 1 || s_.budget_bytes < 0 || !(s_.budget_share > 0 && s_.budget_share <= 1))
		throw std::invalid_argument("server: workers, max_batch and budget must be positive");
	if (sgpu_create(s_.device, &store_) != SGPU_OK)
		throw std::runtime_error("server: no HIP device");
	size_t free_b = 0, total_b = 0;
	if (sgpu_device_memory(store_, &free_b, &total_b) != SGPU_OK)
		throw std::runtime_error("server: device memory query failed");
	budget_ = s_.budget_bytes > 0 ? s_.budget_bytes : s_.budget_share * double(total_b);
	terms_.assign(sf.terms, sf.terms + sf.nterms);
	ci_.assign(sf.ci, sf.ci + sf.nci);
	sf_ = sf;
	sf_.terms = terms_.data();
	sf_.ci = ci_.data();
	t0_ns_ = std::chrono::duration_cast<std::chrono::nanoseconds>(
			std::chrono::steady_clock::now().time_since_epoch()).count();
	for (int k = 0; k < s_.workers; k++)
	{
		std::unique_ptr<Worker> w(new Worker);
		if (sgpu_create(s_.device, &w->ctx) != SGPU_OK ||
				sgpu_set_progress(w->ctx, s_.dock.progress_chunks, s_.dock.progress, s_.dock.progress_user) != SGPU_OK)
			throw std::runtime_error("server: no HIP device");
		workers_.push_back(std::move(w));
	}
	for (auto& w : workers_)
	{
		Worker* wp = w.get();
		w->thread = std::thread([this, wp]() { run(wp); });
	}
}

DockingServer::~DockingServer()
{
	{
		std::lock_guard<std::mutex> lock(mu_);
		stop_ = true;
	}
	cv_work_.notify_all();
	for (auto& w : workers_)
		if (w->thread.joinable())
			w->thread.join();
	for (auto& w : workers_)
	{
		w->receptors.clear();
		sgpu_destroy(w->ctx);
	}
	// the maps' owner goes last: the workers' contexts used them in place
	store_recs_.clear();
	sgpu_destroy(store_);
}

double DockingServer::resident_bytes() const
{
	std::lock_guard<std::mutex> lock(mu_);
	return maps_bytes_;
}

double DockingServer::now() const
{
	const int64_t t = std::chrono::duration_cast<std::chrono::nanoseconds>(
			std::chrono::steady_clock::now().time_since_epoch()).count();
	return double(t - t0_ns_) * 1e-9;
}

int DockingServer::add_receptor(const sgpu_receptor& rec, const sgpu_grid_dims& gd, const std::vector<int>& types)
{
	std::lock_guard<std::mutex> lock(mu_);
	if (next_id_ > 0)
		throw std::logic_error("server: receptors are added before the first job");
	Rec r;
	r.xyz.assign(rec.xyz, rec.xyz + 3 * size_t(rec.natoms));
	r.type.assign(rec.type, rec.type + rec.natoms);
	if (rec.charge)
		r.charge.assign(rec.charge, rec.charge + rec.natoms);
	r.rec = sgpu_receptor{rec.natoms, r.xyz.data(), r.type.data(), r.charge.empty() ? nullptr : r.charge.data()};
	r.gd = gd;
	r.types = types;
	recs_.push_back(std::move(r));
	const Rec& rr = recs_.back();
	const sgpu_receptor rec_view{rr.rec.natoms, rr.xyz.data(), rr.type.data(),
			rr.charge.empty() ? nullptr : rr.charge.data()};
	// the maps, once for the device, in the store context
	std::unique_ptr<GpuDocking> owner(new GpuDocking(s_.dock, sf_, rec_view, rr.gd, store_));
	owner->prepare_types(rr.types);
	const double bytes = double(owner->map_bytes());
	if (maps_bytes_ + bytes > budget_)
		throw std::runtime_error("server: the resident maps exceed the memory budget");
	maps_bytes_ += bytes;
	for (auto& w : workers_)
	{
		// each worker: its tables and receptor atoms, the store's maps in place
		std::unique_ptr<GpuDocking> d(new GpuDocking(s_.dock, sf_, rec_view, rr.gd, w->ctx));
		d->adopt_grid(store_, owner->grid_id(), rr.types);
		w->receptors.push_back(std::move(d));
	}
	store_recs_.push_back(std::move(owner));
	return int(recs_.size()) - 1;
}

int DockingServer::submit(int receptor, const sgpu_model_desc* ligand, int exhaustiveness)
{
	std::lock_guard<std::mutex> lock(mu_);
	if (receptor < 0 || receptor >= int(recs_.size()) || !ligand || exhaustiveness < 1)
		throw std::invalid_argument("server: invalid job");
	std::unique_ptr<Job> j(new Job);
	j->id = next_id_++;
	j->receptor = receptor;
	j->ligand = ligand;
	j->exhaustiveness = exhaustiveness;
	j->t_submit = now();
	queue_.push_back(j.get());
	const int id = j->id;
	jobs_[id] = std::move(j);
	cv_work_.notify_one();
	return id;
}

size_t DockingServer::job_bytes(const Job& j) const
{
	// the device memory a job holds while its launch runs: per chain the
	// state sgpu_run_mc keeps (generator words, the output container of 64
	// poses with confs and heavy-atom coordinates, the model's coordinates,
	// in FP64 the pair energies), per job the model blobs and its FP64 copy
	const size_t n = size_t(j.ligand->natoms);
	const size_t real = s_.dock.precision == SGPU_FP64 ? 8 : 4;
	const size_t conf = 7 + n; // at most one torsion per atom
	const size_t chain = 4096 + 64 * (8 * conf + real * 3 * n) + real * 3 * n + (real == 8 ? 8 * n * n / 2 : 0);
	const size_t chains = size_t(j.exhaustiveness) * size_t(s_.dock.chain_split > 1 ? s_.dock.chain_split : 1);
	return chains * chain + 48 * n * n + (1u << 20);
}

void DockingServer::run(Worker* w)
{
	for (;;)
	{
		std::vector<Job*> batch;
		int receptor = -1;
		double batch_bytes = 0;
		{
			std::unique_lock<std::mutex> lock(mu_);
			cv_work_.wait(lock, [this]() { return stop_ || !queue_.empty(); });
			if (stop_ && queue_.empty())
				return;
			// the receptor of the job that has waited longest, then its queue in
			// arrival order, within the batch size and the device's memory
			// account (resident maps + every launch in flight)
			for (;;)
			{
				if (stop_ && queue_.empty())
					return;
				receptor = queue_.front()->receptor;
				double bytes = 0;
				for (auto it = queue_.begin(); it != queue_.end() && int(batch.size()) < s_.max_batch;)
				{
					Job* j = *it;
					if (j->receptor != receptor)
					{
						++it;
						continue;
					}
					const double b = double(job_bytes(*j));
					if (maps_bytes_ + inflight_bytes_ + bytes + b > budget_)
						break;
					bytes += b;
					batch.push_back(j);
					it = queue_.erase(it);
				}
				if (!batch.empty())
				{
					inflight_bytes_ += bytes;
					peak_inflight_ = std::max(peak_inflight_, inflight_bytes_);
					batch_bytes = bytes;
					break;
				}
				Job* front = queue_.front();
				if (inflight_bytes_ == 0 && maps_bytes_ + double(job_bytes(*front)) > budget_)
				{
					// it can never fit: fail this job alone
					queue_.pop_front();
					front->t_admit = front->t_done = now();
					front->result.error = "job needs " + std::to_string(job_bytes(*front)) +
							" bytes of device memory beside the resident maps, above the server's budget";
					front->result.latency_s = front->t_done - front->t_submit;
					err_ = "job " + std::to_string(front->id) + ": " + front->result.error;
					front->done = true;
					cv_done_.notify_all();
					if (queue_.empty())
						cv_work_.wait(lock, [this]() { return stop_ || !queue_.empty(); });
					continue;
				}
				// wait for a launch in flight to give its memory back
				cv_work_.wait(lock);
				if (queue_.empty())
					cv_work_.wait(lock, [this]() { return stop_ || !queue_.empty(); });
			}
			const double t = now();
			for (Job* j : batch)
				j->t_admit = t;
		}
		std::vector<GpuDocking::DockJob> dj(batch.size());
		std::vector<GpuDocking::DockJob*> dp;
		for (size_t k = 0; k < batch.size(); k++)
		{
			dj[k].m = batch[k]->ligand;
			dj[k].exhaustiveness = batch[k]->exhaustiveness;
			dp.push_back(&dj[k]);
		}
		// a failing launch (an atom overlap, a HIP error, no memory) fails
		// only the jobs that cause it: its jobs are docked again one at a
		// time (a job's search depends on the job alone, so a job docked
		// alone gives the result it would have had in the batch)
		std::vector<std::string> errs(batch.size());
		try
		{
			w->receptors[size_t(receptor)]->dock_batch(dp);
		}
		catch (std::exception& e)
		{
			if (batch.size() == 1)
				errs[0] = e.what();
			else
				for (size_t k = 0; k < batch.size(); k++)
				{
					dj[k].out.clear();
					dj[k].evals = 0;
					std::vector<GpuDocking::DockJob*> one{&dj[k]};
					try
					{
						w->receptors[size_t(receptor)]->dock_batch(one);
					}
					catch (std::exception& e1)
					{
						errs[k] = e1.what();
						dj[k].out.clear();
					}
				}
		}
		{
			std::lock_guard<std::mutex> lock(mu_);
			inflight_bytes_ -= batch_bytes;
			const double t = now();
			for (size_t k = 0; k < batch.size(); k++)
			{
				Job* j = batch[k];
				j->t_done = t;
				j->result.poses = std::move(dj[k].out);
				j->result.evals = dj[k].evals;
				j->result.fp64_error = dj[k].fp64_error;
				j->result.latency_s = j->t_done - j->t_submit;
				j->result.queued_s = j->t_admit - j->t_submit;
				j->result.batch_size = int(batch.size());
				j->result.error = errs[k];
				if (!errs[k].empty())
					err_ = "job " + std::to_string(j->id) + ": " + errs[k];
				j->done = true;
			}
		}
		cv_done_.notify_all();
		cv_work_.notify_all(); // its memory is free again
	}
}

DockingServer::Result DockingServer::wait(int id)
{
	std::unique_lock<std::mutex> lock(mu_);
	auto it = jobs_.find(id);
	if (it == jobs_.end())
		throw std::invalid_argument("server: unknown job");
	Job* j = it->second.get();
	cv_done_.wait(lock, [j]() { return j->done; });
	return j->result;
}

void DockingServer::drain()
{
	std::unique_lock<std::mutex> lock(mu_);
	cv_done_.wait(lock, [this]() {
		for (auto& kv : jobs_)
			if (!kv.second->done)
				return false;
		return true;
	});
}

std::string DockingServer::error() const
{
	std::lock_guard<std::mutex> lock(mu_);
	return err_;
}

} // namespace sgpu_host

// ---------------------------------------------------------------- C wrappers
using namespace sgpu_host;

struct sgpu_server
{
	std::unique_ptr<DockingServer> s;
	std::vector<sgpu_term> terms;
	std::vector<sgpu_ci_term> ci;
	// the message of the last failing call on this server (any thread)
	std::mutex emu;
	std::string err;
	void fail(const std::string& m)
	{
		std::lock_guard<std::mutex> lock(emu);
		err = m;
	}
};

extern "C" {

sgpu_server* sgpu_server_create(const sgpu_server_settings* s)
{
	sgpu_server* sv = new sgpu_server;
	try
	{
		ServerSettings ss;
		ss.device = s->device;
		ss.workers = s->workers;
		ss.max_batch = s->max_batch;
		ss.budget_bytes = s->budget_gb * 1e9; // 0: 75% of the device's memory
		ss.dock.device = s->device;
		if (s->precision != SGPU_FP32 && s->precision != SGPU_FP64 && s->precision != SGPU_MIXED)
			throw std::invalid_argument("precision must be SGPU_FP32, SGPU_FP64 or SGPU_MIXED");
		ss.dock.precision = s->precision;
		ss.dock.seed = s->seed;
		if (s->num_modes < 1 || s->num_modes > 64)
			throw std::invalid_argument("num_modes must be in 1..64");
		ss.dock.num_modes = s->num_modes;
		ss.dock.energy_range = s->energy_range;
		ss.dock.progress_chunks = s->progress_chunks > 0 ? s->progress_chunks : 1;
		ss.dock.progress = s->progress;
		ss.dock.progress_user = s->progress_user;
		if (s->step_fraction < 0 || s->step_fraction > 1)
			throw std::invalid_argument("step_fraction must lie in (0, 1] (0: 1)");
		ss.dock.step_fraction = s->step_fraction > 0 ? s->step_fraction : 1;
		if (s->fast)
		{
			// auto_split's choice when the batch leaves the device idle
			// (src/gpu/gpu_docking.cpp:96-102), fixed: a served job's
			// search must not depend on its batch
			apply_fast(ss.dock);
			ss.dock.chain_split = 32;
			ss.dock.fast_top = 0;
		}
		sv->terms.resize(SGPU_MAX_TERMS);
		sv->ci.resize(SGPU_MAX_CI);
		sgpu_scoring sf;
		const int rc = sgpu_builtin_scoring(s->scoring, s->hydrophobic_scale, sv->terms.data(), sv->ci.data(), &sf);
		if (rc != 0)
			throw std::invalid_argument("unknown scoring function");
		sv->s.reset(new DockingServer(ss, sf));
	}
	catch (std::exception& e)
	{
		sv->fail(e.what());
	}
	return sv;
}

void sgpu_server_free(sgpu_server* sv) { delete sv; }

const char* sgpu_server_error(const sgpu_server* sv)
{
	static thread_local std::string e;
	std::lock_guard<std::mutex> lock(const_cast<sgpu_server*>(sv)->emu);
	e = sv->err;
	return e.c_str();
}

int sgpu_server_add_receptor(sgpu_server* sv, const sgpu_mol* receptor, const sgpu_grid_dims* gd, const int* types,
		int ntypes)
{
	if (!sv->s)
		return -1;
	try
	{
		const sgpu_model_desc* rd = sgpu_mol_desc(receptor);
		return sv->s->add_receptor(sgpu_receptor{rd->natoms, rd->xyz, rd->type, rd->charge}, *gd,
				std::vector<int>(types, types + ntypes));
	}
	catch (std::exception& e)
	{
		sv->fail(e.what());
		return -1;
	}
}

int sgpu_server_submit(sgpu_server* sv, int receptor, const sgpu_mol* ligand, int exhaustiveness)
{
	if (!sv->s)
		return -1;
	try
	{
		return sv->s->submit(receptor, sgpu_mol_desc(ligand), exhaustiveness);
	}
	catch (std::exception& e)
	{
		sv->fail(e.what());
		return -1;
	}
}

int sgpu_server_memory(sgpu_server* sv, double* budget_bytes, double* resident_bytes, double* peak_inflight_bytes)
{
	if (!sv->s)
		return -1;
	if (budget_bytes)
		*budget_bytes = sv->s->budget();
	if (resident_bytes)
		*resident_bytes = sv->s->resident_bytes();
	if (peak_inflight_bytes)
		*peak_inflight_bytes = sv->s->peak_inflight_bytes();
	return 0;
}

int sgpu_server_wait(sgpu_server* sv, int job, sgpu_serve_result* out)
{
	if (!sv->s)
		return -1;
	try
	{
		DockingServer::Result r = sv->s->wait(job);
		out->evals = r.evals;
		out->latency_s = r.latency_s;
		out->queued_s = r.queued_s;
		out->batch_size = r.batch_size;
		out->fp64_error = r.fp64_error;
		if (!r.error.empty())
		{
			out->nposes = 0;
			sv->fail("job " + std::to_string(job) + ": " + r.error);
			return -1;
		}
		out->nposes = int(std::min<size_t>(32, r.poses.size()));
		for (int k = 0; k < out->nposes; k++)
			out->e[k] = r.poses[size_t(k)].e;
		out->evals = r.evals;
		out->latency_s = r.latency_s;
		out->queued_s = r.queued_s;
		out->batch_size = r.batch_size;
		return 0;
	}
	catch (std::exception& e)
	{
		sv->fail(e.what());
		return -1;
	}
}

} // extern "C"
