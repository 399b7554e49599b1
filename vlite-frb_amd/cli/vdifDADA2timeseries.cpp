// vdifDADA2timeseries: VDIF stream -> filterbank -> bandpass -> dedispersion
// at DM 0 and 119.3 with the channel-scatter series -> PRESTO .dat files.
//
// Drop-in for vdifDADA2timeseries.cu:166-763: same flags (-d first, -nchan,
// -nave, -nb, -k hex, -o), NCHANOUT = 3840, 30 s windows with the
// max-delay overlap, seconds-aware gap filling, DM%.1f_<UTC>.dat and
// _SIGMA.dat appended per window, .inf headers and a new set of files at each
// UTC hour.  The PSRDADA shared-memory ring is not supported: the
// VDIF stream is read with --input-file FILE ("-" = stdin).  Extensions:
//   --input-file FILE   VDIF source (file or pipe)
//   --skip-pairs N      frame pairs discarded first (default 256000, initDADA)
//   --rotate-windows K  start new files every K windows instead of each UTC hour
//   --final-inf         also write the .inf headers when the input ends
//   --window S          window length in seconds (default 30, MAXTIME)
//   --dmlist FILE       DM trials, one per line (default 0 and 119.3)
//   --device N
#include <ctime>

#include "cli_common.hpp"

namespace {

struct Args {
    int twonchan = 0, nave = 0, nbundle = 0;
    unsigned key = 0xdada;
    std::string out;
    std::string input;
    long long skip_pairs = 256000;
    int rotate_windows = 0;
    bool final_inf = false;
    int device = 0;
    double window = 30.0;     // s (MAXTIME, TIMESERIES:276-305)
    std::string dmlist;       // one %f per line; default the reference's 0 and 119.3
};

void print_usage() {
    std::fprintf(stderr, "***************************************************************************\n");
    std::fprintf(stderr, "This program reads time series data from a DADA buffer with VDIF format data and\n");
    std::fprintf(stderr, "  generates a filterbank, dedisperses at select DM trials, and outputs time series data.\n\n");
    std::fprintf(stderr, "  Usage:\n./vdifDADA2timeseries -d:\t    use default values (unless other values given)\n");
    std::fprintf(stderr, "\t             -h:            print this help message.\n");
    std::fprintf(stderr, "\t             -k (hex):      hexadecimal key of DADA buffer to read VDIF data from,\n");
    std::fprintf(stderr, "\t             -o (string):   base name of ouput files to write,\n");
    std::fprintf(stderr, "\t             -nave (int):   number of time samples to average over,\n");
    std::fprintf(stderr, "\t             -nchan (int):  number of frequency channels,\n");
    std::fprintf(stderr, "\t             -nb (int):     number of bundles of size (2*2*nchan*nave) to send in 1 memcpy\n");
    std::fprintf(stderr, "\t             --input-file FILE, --skip-pairs N, --rotate-windows K, --final-inf,\n");
    std::fprintf(stderr, "\t             --window S, --dmlist FILE, --device N   (MI355X build extensions)\n");
    std::fprintf(stderr, "\n  DEFAULTS:\n");
    std::fprintf(stderr, "\t             -k     %x\n", 0xdada);
    std::fprintf(stderr, "\t             -o     CRABtestDEDISP_DADA\n");
    std::fprintf(stderr, "\t             -nave  4\n\t             -nchan 6250\n\t             -nb    32\n\n");
}

int parse(int argc, char** argv, Args& a) {
    int i = 1;
    if (cli::eq(argv[i], "-d")) {
        a.twonchan = 6250 * 2;
        a.nave = 4;
        a.nbundle = 32;
        a.key = 0xdada;
        a.out = "testTIMESERIES";
        i++;
    }
    while (i < argc) {
        const char* s = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : ""; };
        if (cli::eq(s, "-nchan")) a.twonchan = 2 * std::atoi(next());
        else if (cli::eq(s, "-nave")) a.nave = std::atoi(next());
        else if (cli::eq(s, "-nb")) a.nbundle = std::atoi(next());
        else if (cli::eq(s, "-k")) std::sscanf(next(), "%x", &a.key);
        else if (cli::eq(s, "-o")) a.out = next();
        else if (cli::eq(s, "--input-file")) a.input = next();
        else if (cli::eq(s, "--skip-pairs")) a.skip_pairs = std::atoll(next());
        else if (cli::eq(s, "--rotate-windows")) a.rotate_windows = std::atoi(next());
        else if (cli::eq(s, "--final-inf")) a.final_inf = true;
        else if (cli::eq(s, "--device")) a.device = std::atoi(next());
        else if (cli::eq(s, "--window")) a.window = std::atof(next());
        else if (cli::eq(s, "--dmlist")) a.dmlist = next();
        else {
            std::fprintf(stderr, "Error! Unknown argument: %s\n\n", s);
            print_usage();
            return 3;
        }
        i++;
    }
    return 0;
}

std::string utc_stamp() {
    time_t t = time(nullptr);
    struct tm u = *gmtime(&t);
    char b[64];
    strftime(b, sizeof b, "%F_%T", &u);
    return b;
}

int utc_hour() {
    time_t t = time(nullptr);
    return gmtime(&t)->tm_hour;
}

struct SeriesFiles {
    std::vector<std::string> base;
    void open_new(const std::vector<float>& dms, const std::string& stamp) {
        base.resize(dms.size());
        for (size_t k = 0; k < dms.size(); k++) {
            char b[256];
            std::snprintf(b, sizeof b, "DM%.1f_%s", dms[k], stamp.c_str());
            base[k] = b;
            for (const char* suf : {".dat", "_SIGMA.dat", ".inf"}) cli::remove_if_exists((base[k] + suf).c_str());
            for (const char* suf : {".dat", "_SIGMA.dat"}) {
                FILE* f = std::fopen((base[k] + suf).c_str(), "wb");
                if (f) std::fclose(f);
            }
        }
    }
    void append(int k, const char* suf, const float* v, size_t n) {
        FILE* f = std::fopen((base[k] + suf).c_str(), "ab");
        if (!f) return;
        std::fwrite(v, sizeof(float), n, f);
        std::fclose(f);
    }
};

bool write_inf(const std::string& base, double mjd, int ntimes, double tsamp, float dm, double fchlo, double bw,
               int nchan, double chanbw) {
    char buf[4096];
    if (vfrb_presto_inf(buf, sizeof buf, base.c_str(), mjd, ntimes, tsamp, dm, fchlo, bw, nchan, chanbw) != VFRB_OK)
        return false;
    FILE* f = std::fopen((base + ".inf").c_str(), "w");
    if (!f) return false;
    std::fputs(buf, f);
    std::fclose(f);
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        print_usage();
        return 1;
    }
    Args a;
    if (parse(argc, argv, a)) return 1;
    if (a.twonchan <= 0 || a.nave <= 0 || a.nbundle <= 0 || a.out.empty()) {
        std::fprintf(stderr, "Error! -nchan, -nave, -nb and -o must be set (use -d for defaults)\n");
        return 1;
    }
    std::string logname = a.out + ".log";
    cli::remove_if_exists(logname.c_str());
    FILE* flog = std::fopen(logname.c_str(), "w");
    if (!flog) {
        std::fprintf(stderr, "Unable to open output log file: %s\n", logname.c_str());
        return 2;
    }
    CLI_CHECK(vfrb_set_device(a.device), 2);
    char devname[256] = "";
    vfrb_device_name(devname, sizeof devname);
    std::fprintf(flog, "Running vdifDADA2timeseries version %s\n", "1.0");
    std::fprintf(flog, "Running on device: %s\n", devname);

    vfrb_params p;
    CLI_CHECK(vfrb_params_derive(a.twonchan, a.nave, a.nbundle, a.window, 0, 3840, &p), 1);
    const int maxloopcnt = p.maxnt / p.nbundle;
    std::fprintf(flog, "NFREQ= %d\nNCHAN= %d\nNCHANOUT= %d\nNAVE= %d\nNBUNDLE= %d\n", p.nfreq, p.nchan, p.nchanout,
                 p.nave, p.nbundle);
    std::fprintf(flog, "MAXNT= %d\nNHALFCYC= %d\nMAXLOOPCNT= %d\n", p.maxnt, p.nhalfcyc, maxloopcnt);
    const double bandw = p.nchanout * p.chanbw;
    const double fchlo = VFRB_LOWFREQ_MHZ + 0.5 * p.chanbw;

    std::fprintf(flog, "DADA buffer at start:\n Connecting to DADA buffer with key: %x ...\n", a.key);
    if (a.input.empty()) {
        std::fprintf(flog, " Could not connect to DADA buffer!\n");
        std::fprintf(stderr, "PSRDADA ring input is not supported by this program; give the VDIF stream with "
                             "--input-file FILE (or - for stdin)\n");
        return -1;
    }
    vfrb_seq* seq = nullptr;
    if (a.input == "-") {
        if (vfrb_seq_open_file(&seq, "/dev/stdin") != VFRB_OK) return 2;
    } else if (vfrb_seq_open_file(&seq, a.input.c_str()) != VFRB_OK) {
        std::fprintf(stderr, "Unable to open input VDIF stream: %s\n", a.input.c_str());
        return 2;
    }
    vfrb_seq_set_log(seq, cli::log_to_file, flog);
    vfrb_seq_set_seconds_check(seq, 1);
    if (vfrb_seq_skip_pairs(seq, a.skip_pairs) != VFRB_OK) {
        std::fprintf(flog, "initDADA returned in error!\n");
        return -1;
    }

    std::vector<float> dms = {0.0f, 119.3f};  // TIMESERIES:441-442
    if (!a.dmlist.empty()) {
        dms.clear();
        FILE* fd = std::fopen(a.dmlist.c_str(), "r");
        if (!fd) {
            std::fprintf(stderr, "Unable to open DM list: %s\n", a.dmlist.c_str());
            return 2;
        }
        float v;
        while (std::fscanf(fd, "%f", &v) == 1) dms.push_back(v);
        std::fclose(fd);
        if (dms.empty()) {
            std::fprintf(stderr, "Error! empty DM list %s\n", a.dmlist.c_str());
            return 1;
        }
    }
    const int ndm = (int)dms.size();
    vfrb_fb_plan* fbp = nullptr;
    if (vfrb_fb_plan_create(&fbp, p.twonchan, p.nave, p.nskip, p.nchanout) != VFRB_OK) {
        std::fprintf(flog, "Error creating FFT plan!\n%s\n", vfrb_last_error());
        return 1;
    }
    vfrb_search_desc sd{};
    sd.nchans = p.nchanout;
    sd.tsamp = p.tchan;
    sd.fch1 = p.fch1;
    sd.chanbw = p.chanbw;
    sd.ndm = ndm;
    sd.dms = dms.data();
    sd.max_rows = p.maxnt;
    sd.storage = VFRB_STORE_F32;
    sd.tile_config = -1;
    vfrb_search_plan* sp = nullptr;
    CLI_CHECK(vfrb_search_plan_create(&sp, &sd), 1);
    const int max_delay = vfrb_search_max_delay(sp);
    const int ntime = p.maxnt - max_delay;
    std::fprintf(flog, "dm_count= %d  max_delay= %d\n", ndm, max_delay);
    if (ntime <= 0) {
        std::fprintf(flog, "ERROR: window of %d rows is not longer than max_delay %d\n", p.maxnt, max_delay);
        return 1;
    }
    std::fprintf(flog, "NTIME_OUTDD= %d\n", ntime);

    const size_t nloop = (size_t)p.nloop;
    const size_t fb_bytes = (size_t)p.maxnt * p.nchanout * sizeof(float);
    const size_t tail_bytes = (size_t)max_delay * p.nchanout * sizeof(float);
    uint8_t *h0, *h1, *d0, *d1;
    float *dfb, *dtail, *dout, *dsig;
    CLI_CHECK(vfrb_host_alloc((void**)&h0, nloop), 2);
    CLI_CHECK(vfrb_host_alloc((void**)&h1, nloop), 2);
    CLI_CHECK(vfrb_malloc((void**)&d0, nloop), 2);
    CLI_CHECK(vfrb_malloc((void**)&d1, nloop), 2);
    CLI_CHECK(vfrb_malloc((void**)&dfb, fb_bytes), 2);
    CLI_CHECK(vfrb_memset(dfb, 0, fb_bytes, nullptr), 2);
    CLI_CHECK(vfrb_malloc((void**)&dtail, tail_bytes ? tail_bytes : 4), 2);
    CLI_CHECK(vfrb_malloc((void**)&dout, (size_t)ndm * ntime * sizeof(float)), 2);
    CLI_CHECK(vfrb_malloc((void**)&dsig, (size_t)ndm * ntime * sizeof(float)), 2);
    std::vector<float> hout((size_t)ndm * ntime), hsig((size_t)ndm * ntime);

    std::fprintf(flog, "initializing pol 0 & 1 single frame buffers...\n");
    vfrb_status ist = vfrb_seq_init(seq);
    if (ist != VFRB_OK) {
        std::fprintf(flog, "initframe01 returned in error code %d!\n", ist == VFRB_ERR_EOF ? 1 : 2);
        return 2;
    }
    uint8_t hdr[32];
    vfrb_seq_header(seq, hdr);
    const double tstart = vfrb_vdif_dmjd(hdr, VFRB_FRAMESPERSEC);
    const double tsamp = p.nave * p.twonchan / VFRB_SAMPLERATE;
    std::fprintf(flog, "tstart= %.15lf   tsamp= %le\n", tstart, tsamp);

    double mjdstart = tstart;
    int ntimes = 0;
    int hrnext = (utc_hour() + 1) % 24;
    std::string stamp = utc_stamp();
    SeriesFiles files;
    files.open_new(dms, stamp);
    std::fclose(flog);
    logname = a.out + "_" + stamp + ".log";
    flog = std::fopen(logname.c_str(), "w");
    if (!flog) {
        std::fprintf(stderr, "Unable to open output log file: %s\n", logname.c_str());
        return 2;
    }
    vfrb_seq_set_log(seq, cli::log_to_file, flog);
    std::fprintf(flog, "begin main loop...\n");

    int loopcnt = 0, q = 0;
    for (;;) {
        if (vfrb_seq_fill(seq, (int32_t)nloop, h0, h1) != VFRB_OK) break;
        CLI_CHECK(vfrb_memcpy_h2d(d0, h0, nloop, nullptr), 2);
        CLI_CHECK(vfrb_memcpy_h2d(d1, h1, nloop, nullptr), 2);
        CLI_CHECK(vfrb_fb_execute(fbp, d0, d1, p.nbundle, dfb, 0, (int64_t)loopcnt * p.nbundle, nullptr), 1);
        loopcnt++;
        if (loopcnt != maxloopcnt) continue;
        CLI_CHECK(vfrb_memcpy_d2d(dtail, dfb + (size_t)ntime * p.nchanout, tail_bytes, nullptr), 2);
        std::fprintf(flog, "calling dedisperseSTATS...\n");
        CLI_CHECK(vfrb_search_load(sp, dfb, 0, 0, p.maxnt, 1, nullptr), 1);
        CLI_CHECK(vfrb_search_dedisperse_stats(sp, p.maxnt, dout, dsig, ntime, nullptr), 1);
        CLI_CHECK(vfrb_memcpy_d2h(hout.data(), dout, hout.size() * sizeof(float), nullptr), 2);
        CLI_CHECK(vfrb_memcpy_d2h(hsig.data(), dsig, hsig.size() * sizeof(float), nullptr), 2);
        for (int k = 0; k < ndm; k++) {
            files.append(k, ".dat", hout.data() + (size_t)k * ntime, ntime);
            files.append(k, "_SIGMA.dat", hsig.data() + (size_t)k * ntime, ntime);
        }
        ntimes += ntime;
        CLI_CHECK(vfrb_memcpy_d2d(dfb, dtail, tail_bytes, nullptr), 2);
        loopcnt = max_delay / p.nbundle;
        q++;
        const bool rollover = a.rotate_windows > 0 ? (q % a.rotate_windows == 0) : (utc_hour() == hrnext);
        if (rollover) {
            for (int k = 0; k < ndm; k++)
                write_inf(files.base[k], mjdstart, ntimes, tsamp, dms[k], fchlo, bandw, p.nchanout, p.chanbw);
            stamp = utc_stamp();
            if (a.rotate_windows > 0) stamp += "_w" + std::to_string(q);
            files.open_new(dms, stamp);
            std::fprintf(flog, "Log file closed at %s\n", stamp.c_str());
            std::fclose(flog);
            logname = a.out + "_" + stamp + ".log";
            flog = std::fopen(logname.c_str(), "w");
            if (!flog) {
                std::fprintf(stderr, "Unable to open output log file: %s\n", logname.c_str());
                return 2;
            }
            vfrb_seq_set_log(seq, cli::log_to_file, flog);
            std::fprintf(flog, "New log file opened at %s\n", stamp.c_str());
            mjdstart = mjdstart + ((double)ntimes * tsamp / 86400.0);
            ntimes = 0;
            hrnext = (hrnext + 1) % 24;
        }
        std::fflush(flog);
    }
    if (a.final_inf)
        for (int k = 0; k < ndm; k++)
            write_inf(files.base[k], mjdstart, ntimes, tsamp, dms[k], fchlo, bandw, p.nchanout, p.chanbw);
    vfrb_seq_close(seq);
    vfrb_fb_plan_destroy(fbp);
    vfrb_search_plan_destroy(sp);
    std::fprintf(flog, "\nfinished!\n");
    std::fclose(flog);
    return 0;
}
