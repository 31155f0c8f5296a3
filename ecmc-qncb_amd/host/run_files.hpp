// run_files.hpp -- the run folder run_dir/run_name and the files in it, in the
// reference's formats (src/io/io.cpp:20-138 and its ILDG and state paths):
//   run_name                       ILDG configuration (the resume point)
//   run_name_params.txt            parameter blocks and "Saved shift" marks
//   run_name_plaquette.txt         one <P> per line, 10 decimals
//   run_name_topo.txt              "t Q E" per flow measurement, 10 decimals
//   run_name_nbevent.txt           "events lifts reversals lambda" (event chains)
//   run_name_seed/run_name_seed<r>.txt   std::mt19937_64 text state of rank r
//   run_name_state/run_name_state<r>.txt event-chain state of rank r
// Observable files are appended to; a run resumes when the configuration and
// rank 0's generator state both exist.
#pragma once
#include <filesystem>
#include <random>
#include <string>
#include <vector>

namespace qcb_drv {

class RunFiles {
public:
    RunFiles(const std::string& run_dir, const std::string& run_name);

    std::filesystem::path folder() const { return folder_; }
    std::filesystem::path configuration() const { return folder_ / name_; }
    std::filesystem::path generator_state(int rank) const;
    std::filesystem::path chain_state(int rank) const;
    std::filesystem::path chain_state_folder() const { return folder_ / (name_ + "_state"); }
    bool resumable() const;

    // Appends text to run_name<suffix> (creating the folder); when `what` is
    // given, reports "<what> <path>" on stdout.  Returns false on failure,
    // which is reported and otherwise ignored, as the reference does.
    bool append(const std::string& suffix, const std::string& text, const char* what = nullptr) const;
    void append_plaquettes(const std::vector<double>& p) const;
    void append_flow_series(const std::vector<double>& tQE) const;   // (t, Q, E) triples
    void log_line(const std::string& line) const { append("_params.txt", line + "\n"); }

    void save_generator(const std::mt19937_64& g, int rank) const;
    bool load_generator(std::mt19937_64& g, int rank) const;   // false (and a message) when absent

private:
    std::filesystem::path folder_;
    std::string name_;
    bool make_folder(const std::filesystem::path& p) const;
};

}  // namespace qcb_drv
