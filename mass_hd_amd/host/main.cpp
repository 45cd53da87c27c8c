// main.cpp -- driver of the host program (src/main.cpp, src/options.cpp,
// src/write.cpp of the reference): option passes, model set-up, the time
// loop with its frame and check cadence, and the JSON frame / parameter
// files.
#include <chrono>
#include <exception>
#include <fstream>
#include <memory>
#include <thread>
#include <iomanip>
#include <iostream>
#include <random>
#include <sstream>
#include <sys/stat.h>

#include "error_msg.hpp"
#include "models.hpp"
#include "options.hpp"
#include "serialization.hpp"

// ---------------------------------------------------------------- globals
unsigned verbose = 2;
ModelPtr model;
std::string model_name;
bool compress = false, compress_full = false;
std::string runname;
std::string output_dir;
std::string inputname;
bool no_write = false, no_warning = false, stop_at_warning = false;
bool force_delete = false, list_models = false;
unsigned pad = 0;
unsigned LX = 0, LY = 0, nthreads = 0, nsteps = 0, ninfo = 0, nstart = 0, nsubsteps = 1, BC = 0;
unsigned seed = 0;
double time_step = 1;
std::chrono::duration<double> write_duration{0};
static VariablesMap vm;
static const unsigned width = 50;

// ---------------------------------------------------------------- output
static void zip_file(const std::string& archive, const std::string& file) {
  const int ret = std::system(inline_str("zip -jm ", archive, " ", file, " > /dev/null 2>&1").c_str());
  if (ret != 0) throw error_msg("zip non-zero return value ", ret, ".");
}

// Frames are formatted and written by a thread of their own while the GPU
// runs the next steps (Step only queues work): the values are downloaded --
// the state at t -- and checked for NaN in the step loop, so a NaN still
// stops the run before the frame's file exists; one frame is in flight at a
// time, and an error of the writer surfaces at the next frame or the end.
static std::thread frame_writer;
static std::exception_ptr frame_writer_error;

static void JoinFrameWriter() {
  if (frame_writer.joinable()) frame_writer.join();
  if (frame_writer_error) {
    std::exception_ptr e = frame_writer_error;
    frame_writer_error = nullptr;
    std::rethrow_exception(e);
  }
}

static void WriteFrame(unsigned t) {
  // src/write.cpp:28-61
  const std::string oname = inline_str(output_dir, "frame", t, ".json");
  auto values = std::make_shared<frame_values>();
  model.ser->serialize_frame(*values);
  if (values->has_nan()) throw error_msg("nan found while writing file.");
  JoinFrameWriter();
  frame_writer = std::thread([values, oname]() {
    try {
      // (the values hold no NaN, so the file is written as the lists are
      // formatted)
      std::ofstream ofs(oname.c_str(), std::ios::out);
      piece_stream buffer;
      buffer.pieces().stream_to(ofs);
      {
        oarchive ar(buffer, "frame", 1);
        values->emit(ar);
      }
      buffer.pieces().finish();
      ofs.close();
      if (compress) zip_file(oname + ".zip", oname);
      if (compress_full) zip_file(runname + ".zip", oname);
    } catch (...) {
      frame_writer_error = std::current_exception();
    }
  });
}

static void WriteParams() {
  // src/write.cpp:63-106
  const std::string oname = inline_str(output_dir, "parameters.json");
  std::stringstream buffer;
  {
    oarchive ar(buffer, "parameters", 1);
    ar.add("LX", LX).add("LY", LY).add("BC", BC).add("nsteps", nsteps).add("nsubsteps", nsubsteps)
        .add("ninfo", ninfo).add("nstart", nstart).add("model_name", model_name);
    model.ser->serialize_params(ar);
    if (ar.bad_value()) throw error_msg("nan found while writing file.");
  }
  std::ofstream ofs(oname.c_str(), std::ios::out);
  ofs << buffer.rdbuf();
  ofs.close();
  if (compress) zip_file(oname + ".zip", oname);
  if (compress_full) zip_file(runname + ".zip", oname);
}

static bool file_exists(const std::string& f) {
  std::ifstream in(f);
  return in.good();
}

static void CreateOutputDir() {
  // src/write.cpp:175-195
  if (compress_full) {
    std::random_device rd;
    output_dir = inline_str("/tmp/mass_hip_", std::hash<std::string>()(inline_str(runname, rd())), "/");
  } else {
    output_dir = runname + (runname.back() == '/' ? "" : "/");
  }
  const int ret = std::system(inline_str("mkdir -p '", output_dir, "'").c_str());
  if (ret) throw error_msg("can not create output directory, mkdir returned ", ret, ".");
}

static void ClearOutput() {
  // src/write.cpp:108-173
  const std::string probe =
      compress_full ? runname + ".zip" : output_dir + "parameters" + (compress ? ".json.zip" : ".json");
  if (!file_exists(probe)) return;
  if (!force_delete) {
    char answ = 0;
    std::cout << "remove output files in '" << (compress_full ? probe : output_dir) << "'? ";
    std::cin >> answ;
    if (answ != 'y' && answ != 'Y')
      throw error_msg("output files already exist in '", (compress_full ? probe : output_dir),
                      "', please provide a different name.");
  }
  const std::string cmd =
      compress_full ? inline_str("rm -f '", probe, "'")
                    : inline_str("rm -f '", output_dir, "'parameters.json '", output_dir,
                                 "'parameters.json.zip '", output_dir, "'frame*.json '",
                                 output_dir, "'frame*.json.zip");
  const int ret = std::system(cmd.c_str());
  if (ret) throw error_msg("rm returned non-zero value ", ret, ".");
}

// ---------------------------------------------------------------- options
static void ParseProgramOptions(int argc, char** argv) {
  // src/options.cpp:38-204
  OptionGroup generic{"Generic options", {}};
  generic.add_switch("help", &list_models, "produce help message").short_name('h');
  generic.opts.back().target = nullptr;  // handled through vm.count("help")
  generic.add_uint("verbose", &verbose, "verbosity level (0=none, 1=little, 2=normal, 3=debug)")
      .short_name('v').implicit_value("2").default_value("2")
      .add_string("input", &inputname, "input file").short_name('i')
      .add_switch("force-delete", &force_delete, "force deletion of existing output file").short_name('f')
      .add_uint("threads", &nthreads, "number of host threads (accepted for compatibility; the step runs on the GPU)")
      .short_name('t').default_value("0").implicit_value("1")
      .add_switch("list", &list_models, "list available models").short_name('l')
      .add_switch("compress", &compress, "compress individual files using zip").short_name('c')
      .add_switch("compress-full", &compress_full, "compress full output using zip (might be slow)")
      .add_switch("no-write", &no_write, "disable file output (for testing purposes)");
  OptionGroup config{"Program options", {}};
  config.add_string("output", &runname, "output name (if compression is on, then .zip is added automatically)")
      .short_name('o')
      .add_string("model", &model_name, "model name (type --list for a list)").short_name('m')
      .add_uint("LX", &LX, "# of nodes in the x direction")
      .add_uint("LY", &LY, "# of nodes in the y direction")
      .add_uint("nsteps", &nsteps, "iterate this many steps in total")
      .add_uint("nsubsteps", &nsubsteps, "subdivision of a single time step").default_value("1")
      .add_uint("ninfo", &ninfo, "save frame every so many steps")
      .add_uint("seed", &seed, "set seed for random number generation (random if unset)")
      .add_switch("no-warning", &no_warning, "disable model specific runtime warnings")
      .add_switch("stop-at-warning", &stop_at_warning, "runtime warnings interrupt the algorithm")
      .add_uint("nstart", &nstart, "time at which to start the output").default_value("0")
      .add_uint("bc", &BC, "boundary conditions flag (0=pbc)").default_value("0");

  // the help switch writes nowhere; give it a dummy target for notify()
  static bool help_flag = false;
  generic.opts[0].target = &help_flag;

  std::vector<const OptionGroup*> cmd = {&generic, &config};
  std::vector<const OptionGroup*> file = {&config};

  vm.clear();
  parse_command_line(argc, argv, cmd, true, vm);
  notify(cmd, vm);
  if (vm.count("help") && model_name.empty() && inputname.empty()) {
    std::cout << describe(cmd) << "\nYou can see model-specific options by specifying a model "
                                  "using '-m' or '--model'\n";
    std::exit(0);
  }
  if (!inputname.empty()) {
    parse_config_file(inputname, file, true, vm);
    notify(cmd, vm);
  }
  if (vm.count("help") && model_name.empty()) {
    std::cout << describe(cmd) << "\n";
    std::exit(0);
  }
  if (list_models) {
    std::cout << "Available models\n" << std::string(width, '=') << "\n" << ListModels() << "\n";
    std::exit(0);
  }
  if (!vm.count("seed")) {
    std::random_device rd;
    seed = rd();
  }
  if (model_name.empty())
    throw error_msg("please specify model. Type --list or -l for a list of available models.");
  model = NewModel(model_name, LX, LY, BC);
  option_list model_options = model->GetOptions();
  static std::vector<OptionGroup> keep;
  keep = model_options;
  for (auto& g : keep) {
    cmd.push_back(&g);
    file.push_back(&g);
  }
  parse_command_line(argc, argv, cmd, false, vm);
  notify(cmd, vm);
  if (vm.count("help")) {
    std::cout << describe(cmd) << "\n";
    std::exit(0);
  }
  if (inputname.empty()) throw error_msg("please provide an input file / type -h for help.");
  parse_config_file(inputname, file, false, vm);
  notify(cmd, vm);
  if (compress_full) compress = false;
  if (!vm.count("output")) runname = compress_full ? "output" : "./";
}

static void ProcessProgramOptions() {
  // src/options.cpp:206-216
  if (!vm.count("nsteps")) throw error_msg("please specify nsteps.");
  if (!vm.count("ninfo") || ninfo == 0) throw error_msg("please specify ninfo > 0.");
  if (!vm.count("LX") || !vm.count("LY")) throw error_msg("please specify LX and LY.");
  pad = (unsigned)inline_str(nsteps).length();
  time_step = 1. / nsubsteps;
  if (nstart % ninfo) nstart = (1u + nstart / ninfo) * ninfo;
}

static void PrintParameters() {
  if (!verbose) return;
  std::cout << "Run parameters\n" << std::string(width, '=') << "\n";
  for (auto& kv : vm.all())
    std::cout << std::left << std::setw(width / 2) << kv.first << std::right << std::setw(width / 2)
              << kv.second.first << "\n";
}

// ---------------------------------------------------------------- run
static void Algorithm() {
  // src/main.cpp:132-197
  for (unsigned t = 0; t < nsteps; ++t) {
    if (t % ninfo == 0) {
      if (!no_write && t >= nstart) {
        const auto start = std::chrono::steady_clock::now();
        WriteFrame(t);
        write_duration += std::chrono::steady_clock::now() - start;
      }
      if (verbose > 1) std::cout << '\n';
      if (verbose)
        std::cout << "timesteps t = " << std::setw(pad) << std::setfill(' ') << std::right << t
                  << " to " << std::setw(pad) << std::setfill(' ') << std::right << t + ninfo
                  << std::endl;
      if (verbose > 1) std::cout << std::string(width, '-') << std::endl;
    }
    for (unsigned s = 0; s < nsubsteps; ++s) model->Step();
    if (t % ninfo == 0 && !no_warning) {
      try {
        if (verbose > 1) model->RuntimeStats();
        model->RuntimeChecks();
      } catch (const warning_msg& e) {
        if (stop_at_warning) throw;
        if (verbose) std::cerr << "warning: " << e.what() << std::endl;
      }
    }
  }
  if (!no_write && nsteps >= nstart) WriteFrame(nsteps);
  const auto start = std::chrono::steady_clock::now();
  JoinFrameWriter();  // the last frame is on disk when the run ends
  write_duration += std::chrono::steady_clock::now() - start;
}

int main(int argc, char** argv) {
  std::cout << "\n   mass (MI355X)  --  lyotropic active nematic on the GPU\n" << std::endl;
  int rc = 0;
  try {
    if (argc < 2) throw error_msg("no argument provided. Type -h for help.");
    DeclareModels();
    ParseProgramOptions(argc, argv);
    ProcessProgramOptions();
    if (runname.empty()) throw error_msg("please specify a file path for this run.");
    PrintParameters();

    if (verbose) std::cout << "\nInitialization\n" << std::string(width, '=') << std::endl;
    if (no_write && verbose) std::cout << "warning: output is not enabled." << std::endl;
    if (verbose) std::cout << "model initialization ..." << std::flush;
    model->Initialize();
    if (verbose) std::cout << " done" << std::endl;
    if (verbose) std::cout << "system initialisation ..." << std::flush;
    model->Configure();
    if (verbose) std::cout << " done" << std::endl;
    if (!no_write) {
      if (verbose) std::cout << "create output directory  ...";
      CreateOutputDir();
      if (verbose) std::cout << " done" << std::endl;
      ClearOutput();
      if (verbose) std::cout << "write parameters ...";
      WriteParams();
      if (verbose) std::cout << " done" << std::endl;
    }
    if (verbose) std::cout << "preparation ... " << std::flush;
    model->Pre();
    if (verbose) std::cout << " done" << std::endl;
    if (verbose) std::cout << "\nRun\n" << std::string(width, '=') << "\n\n";
    model->PreRunStats();
    const auto start = std::chrono::steady_clock::now();
    Algorithm();
    const auto duration = std::chrono::steady_clock::now() - start;
    if (verbose) std::cout << "post-processing ... " << std::flush;
    model->Post();
    if (verbose) std::cout << "done" << std::endl;
    if (verbose) {
      const double secs = std::chrono::duration<double>(duration).count();
      std::cout << "\nStatistics\n" << std::string(width, '=') << "\n";
      std::cout << "Total run time :                    "
                << std::chrono::duration_cast<std::chrono::milliseconds>(duration).count() / 1000.
                << " s\n";
      std::cout << "Total time spent writing output :   "
                << std::chrono::duration_cast<std::chrono::milliseconds>(write_duration).count() /
                       1000.
                << " s\n";
      if (secs > 0)
        std::cout << "Lattice updates per second :        "
                  << (double)LX * LY * nsteps * nsubsteps / secs / 1e6 << " MLUPS (incl. output)\n";
    }
  } catch (const error_msg& e) {
    std::cerr << argv[0] << ": error: " << e.what() << std::endl;
    rc = 1;
  } catch (const std::bad_alloc& ba) {
    std::cerr << argv[0] << ": error initializing memory: " << ba.what() << std::endl;
    rc = 1;
  } catch (const std::exception& e) {
    std::cerr << argv[0] << ": " << e.what() << std::endl;
    rc = 1;
  }
  // a frame still being written when an error stopped the run
  if (frame_writer.joinable()) frame_writer.join();
  return rc;
}
