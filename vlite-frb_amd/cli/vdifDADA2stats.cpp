// vdifDADA2stats: VDIF stream -> filterbank -> bandpass -> per-channel
// statistics per 1 s block, appended to <o>.stats every 60 s window.
//
// Drop-in for vdifDADA2stats.cu:138-454: no reference options besides -h
// (usage, exit 1); defaults TWONCHAN 12500, NAVE 4, NBUNDLE 32, 60 s windows
// of 1 s statistic blocks, no overlap.  The .stats file holds, per window,
// CHANAVE, CHANSIGMA and CHANOUTLIER, each f32 [NSTATBLKS][NCHANOUT].  The
// PSRDADA ring is replaced by --input-file.  Extensions (all additive):
//   --input-file FILE, -o NAME (default vdifDADA2stats_), --nchan N, --nave N,
//   --nb N, --window SEC (60), --statsec SEC (1), --fullband, --device N,
//   --bandpass-always: apply the bandpass at every shape.  By default the
//   program reproduces the reference's output: its
//   subtractbpCAL<<<NCHANOUT,NHALFCYC>>> launch (STATS:248-249,414) exceeds
//   1024 threads per block when NHALFCYC > 1024 (the defaults give 1200) and
//   never runs, so calcstats sees the raw filterbank.
#include "cli_common.hpp"

namespace {

struct Args {
    int twonchan = 12500, nave = 4, nbundle = 32;
    std::string out = "vdifDADA2stats_";
    std::string input;
    double window = 60.0, statsec = 1.0;
    bool fullband = false;
    int device = 0;
    bool bandpass_always = false;
};

void print_usage() {
    std::fprintf(stderr, "***************************************************************************\n");
    std::fprintf(stderr, "This program reads time series data from a DADA buffer with VDIF format data,\n");
    std::fprintf(stderr, "  generates a filterbank and records statistics of the channel power in 1 second intervals.\n\n");
    std::fprintf(stderr, "  Usage:\n./vdifDADA2stats:\n");
    std::fprintf(stderr, "\t             -h:      print this help message.\n");
    std::fprintf(stderr, "\t             --input-file FILE, -o NAME, --nchan N, --nave N, --nb N, --window SEC,\n");
    std::fprintf(stderr, "\t             --statsec SEC, --fullband, --device N, --bandpass-always\n");
    std::fprintf(stderr, "\t             (MI355X build extensions)\n\n");
    std::fprintf(stderr, "  DEFAULTS:\n");
    std::fprintf(stderr, "\t             DADA_INPUT_KEY             = %x\n", 0xdada);
    std::fprintf(stderr, "\t             NAME OF OUTPUT FILE        = vdifDADA2stats.txt\n");
    std::fprintf(stderr, "\t             NUM OF CHANNELS TO AVERAGE = 4\n");
    std::fprintf(stderr, "\t             NUM OF CHANNELS            = 6250\n");
    std::fprintf(stderr, "\t             NUM OF BUNDLES             = 32\n\n");
}

int parse(int argc, char** argv, Args& a) {
    for (int i = 1; i < argc; i++) {
        const char* s = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : ""; };
        if (cli::eq(s, "-h")) {
            print_usage();
            return 3;
        } else if (cli::eq(s, "--input-file")) a.input = next();
        else if (cli::eq(s, "-o")) a.out = next();
        else if (cli::eq(s, "--nchan")) a.twonchan = 2 * std::atoi(next());
        else if (cli::eq(s, "--nave")) a.nave = std::atoi(next());
        else if (cli::eq(s, "--nb")) a.nbundle = std::atoi(next());
        else if (cli::eq(s, "--window")) a.window = std::atof(next());
        else if (cli::eq(s, "--statsec")) a.statsec = std::atof(next());
        else if (cli::eq(s, "--fullband")) a.fullband = true;
        else if (cli::eq(s, "--device")) a.device = std::atoi(next());
        else if (cli::eq(s, "--bandpass-always")) a.bandpass_always = true;
        else {
            std::fprintf(stderr, "Error! Unknown argument: %s\n\n", s);
            print_usage();
            return 3;
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    if (parse(argc, argv, a)) return 1;
    const std::string logname = a.out + ".log", statsname = a.out + ".stats";
    FILE* flog = std::fopen(logname.c_str(), "w");
    if (!flog) {
        std::fprintf(stderr, "Unable to open output log file: %s\n", logname.c_str());
        return 2;
    }
    FILE* fst = std::fopen(statsname.c_str(), "wb");
    if (!fst) {
        std::fprintf(stderr, "Unable to open output STATS file: %s\n", statsname.c_str());
        return 2;
    }
    std::fclose(fst);
    CLI_CHECK(vfrb_set_device(a.device), 2);
    char devname[256] = "";
    vfrb_device_name(devname, sizeof devname);
    std::fprintf(flog, "Running vdifDADA2stats version %s\nRunning on device: %s\n", "1.0", devname);

    // STATS:209-232: MAXNT and the stat blocks from float MAXTIME/TIMESTATS.
    vfrb_params p;
    CLI_CHECK(vfrb_params_derive(a.twonchan, a.nave, a.nbundle, a.window, a.fullband ? 1 : 0, 0, &p), 1);
    const float maxtime = (float)a.window, timestats = (float)a.statsec;
    const double tsamp = p.nave * p.twonchan / VFRB_SAMPLERATE;
    const int nstatblks = (int)(maxtime / timestats);
    const int ncolstatblk = (int)(timestats / tsamp);
    const int nloopwin = p.maxnt / p.nbundle;
    std::fprintf(flog, "NFREQ   = %d\nNCHAN   = %d\nNCHANOUT= %d\nNAVE    = %d\nNBUNDLE = %d\nNLOOP   = %d\n",
                 p.nfreq, p.nchan, p.nchanout, p.nave, p.nbundle, nloopwin);
    const bool bandpass = a.bandpass_always || p.nhalfcyc <= cli::kRefMaxBlockThreads;
    if (!bandpass)
        std::fprintf(flog, "NHALFCYC= %d > %d: statistics of the raw filterbank, as in the reference "
                           "(--bandpass-always applies the bandpass)\n", p.nhalfcyc, cli::kRefMaxBlockThreads);
    if (nstatblks <= 0 || ncolstatblk <= 0 || (int64_t)ncolstatblk * (nstatblks - 1) >= p.maxnt) {
        std::fprintf(stderr, "Error! stat blocks do not fit the window\n");
        return 1;
    }
    std::fprintf(flog, "DADA buffer at start:\n Connecting to DADA buffer with key: %x ...\n", 0xdada);
    if (a.input.empty()) {
        std::fprintf(flog, " Could not connect to DADA buffer!\n");
        std::fprintf(stderr, "PSRDADA ring input is not supported by this program; give the VDIF stream with "
                             "--input-file FILE (or - for stdin)\n");
        return -1;
    }
    vfrb_seq* seq = nullptr;
    if (vfrb_seq_open_file(&seq, a.input == "-" ? "/dev/stdin" : a.input.c_str()) != VFRB_OK) {
        std::fprintf(stderr, "Unable to open input VDIF stream: %s\n", a.input.c_str());
        return 2;
    }
    vfrb_seq_set_log(seq, cli::log_to_file, flog);

    vfrb_fb_plan* fbp = nullptr;
    if (vfrb_fb_plan_create(&fbp, p.twonchan, p.nave, p.nskip, p.nchanout) != VFRB_OK) {
        std::fprintf(flog, "Error creating plan!\n%s\n", vfrb_last_error());
        return 1;
    }
    const size_t nloop = (size_t)p.nloop;
    const size_t fb_bytes = (size_t)p.maxnt * p.nchanout * sizeof(float);
    const size_t blk_bytes = (size_t)nstatblks * p.nchanout * sizeof(float);
    uint8_t *h0, *h1, *d0, *d1;
    float *dfb, *dave, *dsig, *dout;
    CLI_CHECK(vfrb_host_alloc((void**)&h0, nloop), 2);
    CLI_CHECK(vfrb_host_alloc((void**)&h1, nloop), 2);
    CLI_CHECK(vfrb_malloc((void**)&d0, nloop), 2);
    CLI_CHECK(vfrb_malloc((void**)&d1, nloop), 2);
    CLI_CHECK(vfrb_malloc((void**)&dfb, fb_bytes), 2);
    CLI_CHECK(vfrb_memset(dfb, 0, fb_bytes, nullptr), 2);
    CLI_CHECK(vfrb_malloc((void**)&dave, blk_bytes), 2);
    CLI_CHECK(vfrb_malloc((void**)&dsig, blk_bytes), 2);
    CLI_CHECK(vfrb_malloc((void**)&dout, blk_bytes), 2);
    std::vector<float> have(blk_bytes / 4), hsig(blk_bytes / 4), hout(blk_bytes / 4);

    std::fprintf(flog, "initializing pol 0 & 1 single frame buffers...\n");
    if (vfrb_seq_init(seq) != VFRB_OK) {
        std::fprintf(flog, "initframe01 returned in error!\n");
        return 2;
    }
    int loopcnt = 0;
    std::fprintf(flog, "begin main loop...\n");
    for (;;) {
        if (vfrb_seq_fill(seq, (int32_t)nloop, h0, h1) != VFRB_OK) break;
        CLI_CHECK(vfrb_memcpy_h2d(d0, h0, nloop, nullptr), 2);
        CLI_CHECK(vfrb_memcpy_h2d(d1, h1, nloop, nullptr), 2);
        CLI_CHECK(vfrb_fb_execute(fbp, d0, d1, p.nbundle, dfb, 0, (int64_t)loopcnt * p.nbundle, nullptr), 1);
        loopcnt++;
        if (loopcnt == nloopwin) {
            // STATS:414-416: subtractbpCAL then calcstats, as one pass
            if (bandpass)
                CLI_CHECK(vfrb_bandpass_chan_stats(dfb, p.maxnt, p.nchanout, ncolstatblk, nstatblks, dave, dsig,
                                                   dout, nullptr), 1);
            else
                CLI_CHECK(vfrb_chan_stats(dfb, p.maxnt, p.nchanout, ncolstatblk, nstatblks, dave, dsig, dout,
                                          nullptr), 1);
            CLI_CHECK(vfrb_memcpy_d2h(have.data(), dave, blk_bytes, nullptr), 2);
            CLI_CHECK(vfrb_memcpy_d2h(hsig.data(), dsig, blk_bytes, nullptr), 2);
            CLI_CHECK(vfrb_memcpy_d2h(hout.data(), dout, blk_bytes, nullptr), 2);
            FILE* f = std::fopen(statsname.c_str(), "ab");
            if (f) {
                std::fwrite(have.data(), blk_bytes, 1, f);
                std::fwrite(hsig.data(), blk_bytes, 1, f);
                std::fwrite(hout.data(), blk_bytes, 1, f);
                std::fclose(f);
            }
            loopcnt = 0;
        }
    }
    std::fprintf(flog, "Looped %d times\n\n", loopcnt);
    vfrb_seq_close(seq);
    vfrb_fb_plan_destroy(fbp);
    std::fprintf(flog, "Processed %f [sec] of data\n", loopcnt * p.tbundle);
    std::fprintf(flog, "\nfinished!\n");
    std::fclose(flog);
    return 0;
}
