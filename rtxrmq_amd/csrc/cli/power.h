// power.h -- GPU and CPU power samplers for --save-power.
//
// The reference polls NVML every 100 ms on a pthread and writes one row per
// sample: timestep, power (W), accumulated energy (J), average power, dt,
// accumulated time (common/nvmlPower.cpp:37-84).  Here the amdgpu driver's
// hwmon sensor (power1_average or power1_input, microwatts) of the card whose
// PCI bus id matches the HIP device is read instead; rows keep the same
// columns so the reference's RMQ/J scripts read them unchanged.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <dirent.h>
#include <string>
#include <thread>
#include <vector>

class PowerSampler {
   public:
    ~PowerSampler() { stop(); }

    // Starts sampling every period_ms into `file` (default ../results/power-<alg>.dat).
    bool start(int dev, const std::string &file, const char *alg, int period_ms = 100) {
        stop();
        sensor_ = find_sensor(dev);
        if (sensor_.empty()) {
            fprintf(stderr, "--save-power: no readable amdgpu power sensor for device %d; skipping\n", dev);
            return false;
        }
        std::string path = file;
        if (path.empty()) {
            path = "../results/power-";
            for (const char *c = alg; *c; ++c) path += std::isalnum((unsigned char)*c) ? *c : '_';
            path += ".dat";
        }
        out_ = fopen(path.c_str(), "w+");
        if (!out_) {
            fprintf(stderr, "--save-power: cannot open %s; skipping\n", path.c_str());
            return false;
        }
        fprintf(out_, "%-15s %-15s %-15s %-15s %-15s %-15s\n", "#timestep", "power", "acc-energy", "avg-power", "dt",
                "acc-time");
        running_ = true;
        thread_ = std::thread([this, period_ms] { loop(period_ms); });
        return true;
    }

    void stop() {
        if (!running_) return;
        running_ = false;
        if (thread_.joinable()) thread_.join();
        if (out_) fclose(out_);
        out_ = nullptr;
        printf("power: %.3f J over %.3f s, average %.2f W\n", energy_j_, time_s_, time_s_ > 0 ? energy_j_ / time_s_ : 0.0);
    }

    // Energy accumulated between start() and stop(), joules (0 without a sensor).
    double energy_j() const { return energy_j_; }

   private:
    static double read_watts(const std::string &sensor) {
        FILE *f = fopen(sensor.c_str(), "r");
        if (!f) return -1;
        long long uw = -1;
        if (fscanf(f, "%lld", &uw) != 1) uw = -1;
        fclose(f);
        return uw < 0 ? -1 : uw / 1e6;
    }

    static std::string find_sensor(int dev) {
        char bus[64] = {0};
        if (hipDeviceGetPCIBusId(bus, sizeof(bus), dev) != hipSuccess) return "";
        for (char *c = bus; *c; ++c) *c = (char)std::tolower((unsigned char)*c);
        const std::string base = std::string("/sys/bus/pci/devices/") + bus + "/hwmon";
        DIR *d = opendir(base.c_str());
        if (!d) return "";
        std::string found;
        while (dirent *e = readdir(d)) {
            if (std::string(e->d_name).rfind("hwmon", 0) != 0) continue;
            for (const char *name : {"power1_average", "power1_input"}) {
                const std::string p = base + "/" + e->d_name + "/" + name;
                if (read_watts(p) >= 0) { found = p; break; }
            }
            if (!found.empty()) break;
        }
        closedir(d);
        return found;
    }

    void loop(int period_ms) {
        using clk = std::chrono::steady_clock;
        auto t1 = clk::now();
        int step = 0;
        while (running_) {
            std::this_thread::sleep_for(std::chrono::milliseconds(period_ms));
            const auto t2 = clk::now();
            const double dt = std::chrono::duration<double>(t2 - t1).count();
            t1 = t2;
            const double w = read_watts(sensor_);
            if (w < 0) continue;
            ++step;
            time_s_ += dt;
            energy_j_ += w * dt;
            fprintf(out_, "%-15i %-15f %-15f %-15f %-15f %-15f\n", step, w, energy_j_, energy_j_ / time_s_, dt, time_s_);
        }
    }

    std::string sensor_;
    FILE *out_ = nullptr;
    std::atomic<bool> running_{false};
    std::thread thread_;
    double energy_j_ = 0, time_s_ = 0;
};

// CPU package energy for the CPU algorithms, as the reference samples RAPL
// every 500 ms around its CPU methods (common/nvmlPower.cpp:216-274, reading
// MSR_PKG_ENERGY_STATUS through /dev/cpu/N/msr in common/Rapl.cpp).  Here the
// kernel's powercap interface is read instead, which needs no MSR access:
// every /sys/class/powercap/<zone> whose name starts with "package" and whose
// energy_uj is readable (AMD and Intel packages alike), summed over sockets,
// with the counter's wrap at max_energy_range_uj.  Rows keep the reference's
// six columns: timestep, power, accumulated energy, average power, dt,
// accumulated time.  RTXRMQ_POWERCAP overrides the class directory.
class CpuEnergySampler {
   public:
    ~CpuEnergySampler() { stop(); }

    bool start(const std::string &file, const char *alg, int period_ms = 500) {
        stop();
        zones_ = find_zones();
        if (zones_.empty()) {
            fprintf(stderr, "--save-power: no readable powercap package energy counter; skipping\n");
            return false;
        }
        std::string path = file;
        if (path.empty()) {
            path = "../results/power-";
            for (const char *c = alg; *c; ++c) path += std::isalnum((unsigned char)*c) ? *c : '_';
            path += ".dat";
        }
        out_ = fopen(path.c_str(), "w+");
        if (!out_) {
            fprintf(stderr, "--save-power: cannot open %s; skipping\n", path.c_str());
            return false;
        }
        fprintf(out_, "%-15s %-15s %-15s %-15s %-15s %-15s\n", "#timestep", "power", "acc-energy", "avg-power", "dt",
                "acc-time");
        for (Zone &z : zones_) z.last = read_uj(z.energy);
        running_ = true;
        thread_ = std::thread([this, period_ms] { loop(period_ms); });
        return true;
    }

    void stop() {
        if (!running_) return;
        running_ = false;
        if (thread_.joinable()) thread_.join();
        if (out_) fclose(out_);
        out_ = nullptr;
        printf("CPU power: %.3f J over %.3f s, average %.2f W (%zu package%s)\n", energy_j_, time_s_,
               time_s_ > 0 ? energy_j_ / time_s_ : 0.0, zones_.size(), zones_.size() == 1 ? "" : "s");
    }

    double energy_j() const { return energy_j_; }

   private:
    struct Zone {
        std::string energy;
        long long range = 0;  // max_energy_range_uj: the counter wraps to 0 past it
        long long last = -1;
    };

    static long long read_uj(const std::string &path) {
        FILE *f = fopen(path.c_str(), "r");
        if (!f) return -1;
        long long v = -1;
        if (fscanf(f, "%lld", &v) != 1) v = -1;
        fclose(f);
        return v;
    }

    static std::vector<Zone> find_zones() {
        const char *env = getenv("RTXRMQ_POWERCAP");
        const std::string base = env && *env ? env : "/sys/class/powercap";
        std::vector<Zone> zones;
        DIR *d = opendir(base.c_str());
        if (!d) return zones;
        while (dirent *e = readdir(d)) {
            if (e->d_name[0] == '.') continue;
            const std::string dir = base + "/" + e->d_name;
            char name[64] = {0};
            FILE *f = fopen((dir + "/name").c_str(), "r");
            if (!f) continue;
            const bool ok = fscanf(f, "%63s", name) == 1;
            fclose(f);
            if (!ok || std::string(name).rfind("package", 0) != 0) continue;
            Zone z;
            z.energy = dir + "/energy_uj";
            z.range = read_uj(dir + "/max_energy_range_uj");
            if (read_uj(z.energy) >= 0) zones.push_back(z);
        }
        closedir(d);
        return zones;
    }

    void loop(int period_ms) {
        using clk = std::chrono::steady_clock;
        auto t1 = clk::now();
        int step = 0;
        while (running_) {
            std::this_thread::sleep_for(std::chrono::milliseconds(period_ms));
            const auto t2 = clk::now();
            const double dt = std::chrono::duration<double>(t2 - t1).count();
            t1 = t2;
            double joules = 0;
            for (Zone &z : zones_) {
                const long long now = read_uj(z.energy);
                if (now < 0 || z.last < 0) continue;
                long long d = now - z.last;
                if (d < 0 && z.range > 0) d += z.range;  // wrapped
                if (d > 0) joules += d / 1e6;
                z.last = now;
            }
            ++step;
            time_s_ += dt;
            energy_j_ += joules;
            fprintf(out_, "%-15i %-15f %-15f %-15f %-15f %-15f\n", step, dt > 0 ? joules / dt : 0.0, energy_j_,
                    energy_j_ / time_s_, dt, time_s_);
        }
    }

    std::vector<Zone> zones_;
    FILE *out_ = nullptr;
    std::atomic<bool> running_{false};
    std::thread thread_;
    double energy_j_ = 0, time_s_ = 0;
};
