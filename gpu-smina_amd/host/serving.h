// Docking server: several receptors resident on one GPU, jobs admitted as
// they arrive and batched continuously.
//
// Each worker thread owns a device context (one launch at a time per
// context); every receptor's maps are built once, in the server's store
// context, and every worker uses them in place (sgpu_share_grid), so the
// maps are resident once per device.  A free worker takes the waiting jobs
// of the receptor whose oldest job waited longest, up to a batch size, and
// admits them only while one device-wide account holds them: the resident
// maps plus the chain state and models of every launch in flight, within
// the server's budget (a stated share of the device's memory).  It docks
// them as one launch (GpuDocking::dock_batch semantics per job: search,
// merge, refinement, FP64 rescoring).  With W workers, up to W launches
// are on the device at once, so jobs that arrive while others run start as
// soon as a worker and the memory free up, and their work overlaps the
// running launches.  A job's chains are seeded by the job alone, so its
// result does not depend on what it was batched with.
//
// Each worker's context drives up to four streams.  Streams that share a
// hardware queue run their kernels in FIFO order, so the process needs
// GPU_MAX_HW_QUEUES >= 4 x workers (set before the HIP runtime starts) for
// a short job never to wait behind another worker's long launch.
#ifndef SGPU_SERVING_H
#define SGPU_SERVING_H

#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gpu_docking.h"

namespace sgpu_host
{

struct ServerSettings
{
	int device = 0;
	int workers = 2;            // concurrent launches
	int max_batch = 64;         // ligands per launch
	// device memory for resident maps and in-flight jobs, one account for
	// the device; 0: budget_share of the device's total memory
	double budget_bytes = 0;
	double budget_share = 0.75;
	GpuSettings dock;           // precision, seed, modes, energy range
};

class DockingServer
{
public:
	explicit DockingServer(const ServerSettings& s, const sgpu_scoring& sf);
	~DockingServer();
	DockingServer(const DockingServer&) = delete;
	DockingServer& operator=(const DockingServer&) = delete;

	// a resident receptor with its box and the ligand types its maps cover;
	// returns its index
	int add_receptor(const sgpu_receptor& rec, const sgpu_grid_dims& gd, const std::vector<int>& types);
	// queue a job; returns its id.  The ligand description must stay valid
	// until the job completes.
	int submit(int receptor, const sgpu_model_desc* ligand, int exhaustiveness);
	struct Result
	{
		std::vector<Pose> poses;
		int64_t evals = 0;
		double latency_s = 0; // submit to result
		double queued_s = 0;  // submit to admission
		int batch_size = 0;   // jobs in its launch
		double fp64_error = -1; // DockJob::fp64_error of the job
		std::string error;    // empty: docked; else why this job failed
	};
	// block until job `id` completes
	Result wait(int id);
	// block until every submitted job completes
	void drain();
	std::string error() const;
	// the device-wide memory account: its budget, the resident maps (once
	// per device) and the peak of the launches in flight
	double budget() const { return budget_; }
	double resident_bytes() const;
	double peak_inflight_bytes() const
	{
		std::lock_guard<std::mutex> lock(mu_);
		return peak_inflight_;
	}

private:
	struct Job
	{
		int id;
		int receptor;
		const sgpu_model_desc* ligand;
		int exhaustiveness;
		double t_submit = 0, t_admit = 0, t_done = 0;
		bool done = false;
		Result result;
	};
	struct Worker
	{
		sgpu_ctx* ctx = nullptr;
		std::vector<std::unique_ptr<GpuDocking>> receptors;
		std::thread thread;
	};

	void run(Worker* w);
	double now() const;
	size_t job_bytes(const Job& j) const;

	ServerSettings s_;
	std::vector<sgpu_term> terms_;
	std::vector<sgpu_ci_term> ci_;
	sgpu_scoring sf_;
	struct Rec
	{
		sgpu_receptor rec;
		std::vector<double> xyz;
		std::vector<int> type;
		std::vector<double> charge;
		sgpu_grid_dims gd;
		std::vector<int> types;
	};
	std::vector<Rec> recs_;
	// the store context holds every receptor's maps (destroyed after the
	// workers, whose contexts use them)
	sgpu_ctx* store_ = nullptr;
	std::vector<std::unique_ptr<GpuDocking>> store_recs_;
	double budget_ = 0;
	double maps_bytes_ = 0;
	double inflight_bytes_ = 0;
	double peak_inflight_ = 0;
	std::condition_variable cv_mem_;
	std::vector<std::unique_ptr<Worker>> workers_;
	mutable std::mutex mu_;
	std::condition_variable cv_work_, cv_done_;
	std::deque<Job*> queue_;
	std::map<int, std::unique_ptr<Job>> jobs_;
	int next_id_ = 0;
	bool stop_ = false;
	std::string err_;
	int64_t t0_ns_ = 0;
};

} // namespace sgpu_host

#endif
