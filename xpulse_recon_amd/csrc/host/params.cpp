#include "params.hpp"

#include "../../../include/xpr_hip.h"

#include <fstream>
#include <sstream>
#include <stdexcept>

namespace xpr {

const char *method_name(Method m)
{
    switch (m) {
    case Method::naive: return "naive";
    case Method::method_1a: return "method_1a";
    case Method::method_1b: return "method_1b";
    case Method::method_1c: return "method_1c";
    case Method::method_2: return "method_2";
    case Method::hartmann: return "hartmann";
    case Method::ipr: return "ipr";
    }
    return "?";
}

namespace {

[[noreturn]] void bad(const std::string &key, const std::string &what)
{
    throw std::runtime_error("parameter '" + key + "': " + what);
}

void rd(const Json &o, const char *key, std::string &v, const std::string &sec)
{
    if (const Json *j = o.get(key)) {
        if (j->kind != Json::String) bad(sec + "." + key, "expected a string");
        v = j->str;
    }
}
void rd(const Json &o, const char *key, bool &v, const std::string &sec)
{
    if (const Json *j = o.get(key)) {
        if (j->kind != Json::Bool) bad(sec + "." + key, "expected true or false");
        v = j->b;
    }
}
template <typename T>
void rdn(const Json &o, const char *key, T &v, const std::string &sec)
{
    if (const Json *j = o.get(key)) {
        if (j->kind != Json::Number) bad(sec + "." + key, "expected a number");
        v = T(j->num);
    }
}

Method parse_method(const std::string &s)
{
    for (Method m : {Method::naive, Method::method_1a, Method::method_1b, Method::method_1c,
                     Method::method_2, Method::hartmann, Method::ipr})
        if (s == method_name(m)) return m;
    bad("recon.method", "unknown method '" + s + "'");
}

const Json &section(const Json &root, const char *name, const Json &empty)
{
    const Json *j = root.get(name);
    if (!j || j->kind == Json::Null) return empty;
    if (j->kind != Json::Object) bad(name, "expected an object");
    return *j;
}

} // namespace

Parameters parameters_from_json(const Json &root)
{
    if (root.kind != Json::Object) throw std::runtime_error("parameters must be a JSON object");
    Parameters p;
    const Json empty = Json::object();
    {
        const Json &r = section(root, "recon", empty);
        auto &q = p.recon;
        rd(r, "input", q.input, "recon");
        rd(r, "output", q.output, "recon");
        rd(r, "proj_output", q.proj_output, "recon");
        if (const Json *m = r.get("method")) {
            if (m->kind != Json::String) bad("recon.method", "expected a string");
            q.method = parse_method(m->str);
        }
        rdn(r, "it", q.it, "recon");
        rdn(r, "sit", q.sit, "recon");
        rd(r, "mlog", q.mlog, "recon");
        rdn(r, "weight", q.weight, "recon");
        rdn(r, "weight_factor", q.weight_factor, "recon");
        rd(r, "normalize", q.normalize, "recon");
        rd(r, "images_preproc", q.normalize, "recon");   // README name of the same switch
        rd(r, "simulation", q.simulation, "recon");
        if (q.it < 1) bad("recon.it", "must be >= 1");
        if (q.sit < 1) bad("recon.sit", "must be >= 1");
    }
    {
        const Json &g = section(root, "geometry", empty);
        auto &q = p.geometry;
        rdn(g, "so", q.so, "geometry");
        rdn(g, "sd", q.sd, "geometry");
        rdn(g, "angle", q.angle, "geometry");
        rdn(g, "heli_offset", q.heli_offset, "geometry");
        rdn(g, "heli_step", q.heli_step, "geometry");
        if (const Json *a = g.get("angle_list")) {
            if (a->kind != Json::Array) bad("geometry.angle_list", "expected an array");
            for (const Json &e : a->arr) {
                if (e.kind != Json::Number) bad("geometry.angle_list", "expected numbers");
                q.angle_list.push_back(float(e.num));
            }
        }
    }
    {
        const Json &d = section(root, "detector", empty);
        auto &q = p.detector;
        rdn(d, "px", q.px, "detector");
        rdn(d, "rx", q.rx, "detector");
        rdn(d, "sx", q.sx, "detector");
        rdn(d, "sy", q.sy, "detector");
        rdn(d, "module_number", q.module_number, "detector");
        if (const Json *m = d.get("module_angle_offset")) {
            if (m->kind == Json::Number) q.module_angle_offset = float(m->num);
            else if (m->kind != Json::Null) bad("detector.module_angle_offset", "expected a number");
        }
        rdn(d, "module_center_offset_z", q.module_center_offset_z, "detector");
        rdn(d, "roll", q.roll, "detector");
        rdn(d, "pitch", q.pitch, "detector");
        rdn(d, "yaw", q.yaw, "detector");
        if (const Json *ms = d.get("modules")) {           // README form (README.md:71-77)
            if (ms->kind == Json::Array && !ms->arr.empty() && ms->arr[0].kind == Json::Object) {
                rdn(ms->arr[0], "roll", q.roll, "detector.modules[0]");
                rdn(ms->arr[0], "pitch", q.pitch, "detector.modules[0]");
                rdn(ms->arr[0], "yaw", q.yaw, "detector.modules[0]");
            }
        }
        if (q.module_number < 1) bad("detector.module_number", "must be >= 1");
    }
    {
        const Json &m = section(root, "method2", empty);
        rdn(m, "overlap_fct", p.method2.overlap_fct, "method2");
    }
    {
        const Json &h = section(root, "hartmann", empty);
        rd(h, "devx", p.hartmann.devx, "hartmann");
        rd(h, "devy", p.hartmann.devy, "hartmann");
        rdn(h, "iterations", p.hartmann.iterations, "hartmann");
        rdn(h, "iterations_fct", p.hartmann.iterations_fct, "hartmann");
        rdn(h, "iterations_step", p.hartmann.iterations_step, "hartmann");
        rd(h, "method", p.hartmann.method, "hartmann");
    }
    {
        const Json &i = section(root, "ipr", empty);
        rdn(i, "iterations", p.ipr.iterations, "ipr");
        rdn(i, "iterations_fct", p.ipr.iterations_fct, "ipr");
        rdn(i, "iterations_step", p.ipr.iterations_step, "ipr");
        rdn(i, "energy_kev", p.ipr.energy_kev, "ipr");
    }
    {
        const Json &x = section(root, "xpr", empty);
        auto &q = p.xpr;
        if (const Json *s = x.get("storage")) {
            if (s->kind != Json::String) bad("xpr.storage", "expected a string");
            if (s->str == "f32") q.storage = 0;
            else if (s->str == "f16") q.storage = 1;
            else if (s->str == "bf16") q.storage = 2;
            else bad("xpr.storage", "one of f32, f16, bf16");
        }
        rd(x, "final_fp32", q.final_fp32, "xpr");
        if (const Json *a = x.get("accumulator")) {
            if (a->kind == Json::String) {
                if (a->str != "f32") bad("xpr.accumulator", "32, 64 or \"f32\"");
                q.accumulator = XPR_ACC_F32;
            } else {
                rdn(x, "accumulator", q.accumulator, "xpr");
                if (q.accumulator != 32 && q.accumulator != 64)
                    bad("xpr.accumulator", "32, 64 or \"f32\"");
            }
        }
        if (const Json *s = x.get("interp")) {
            if (s->kind != Json::String) bad("xpr.interp", "expected a string");
            if (s->str == "nearest") q.interp = XPR_INTERP_NEAREST;
            else if (s->str == "bilinear") q.interp = XPR_INTERP_BILINEAR;
            else bad("xpr.interp", "nearest or bilinear");
        }
        rd(x, "jitter", q.jitter, "xpr");
        rdn(x, "jitter_seed", q.jitter_seed, "xpr");
        rd(x, "geometry_fp64", q.geometry_fp64, "xpr");
        rdn(x, "slab_height", q.slab_height, "xpr");
        if (q.slab_height < 0) bad("xpr.slab_height", "must be >= 0 (0: from the device budget)");
        rdn(x, "device_budget_gb", q.device_budget_gb, "xpr");
        rd(x, "layer_store", q.layer_store, "xpr");
        if (q.layer_store != "zstd" && q.layer_store != "ram" && q.layer_store != "device" &&
            q.layer_store != "auto")
            bad("xpr.layer_store", "device, ram, zstd or auto");
        rdn(x, "threads", q.threads, "xpr");
        rd(x, "run_log", q.run_log, "xpr");
        rd(x, "autotune", q.autotune, "xpr");
        rd(x, "narrow_sums", q.narrow_sums, "xpr");
        rd(x, "quiet", q.quiet, "xpr");
        rdn(x, "device", q.device, "xpr");
    }
    // initializeParameters (parameters.hpp:34-48); rx stays in pixels here and
    // is multiplied by px where the geometry is built.
    if (p.method2.overlap_fct > 1.0f) p.method2.overlap_fct = 1.0f;
    if (p.method2.overlap_fct < 0.01f) p.method2.overlap_fct = 0.01f;
    if (p.geometry.heli_step < 1.0f) p.geometry.heli_step = 1.0f;
    return p;
}

Json parameters_to_json(const Parameters &p)
{
    Json root = Json::object();
    Json r = Json::object();
    r.set("input", Json::string(p.recon.input));
    r.set("output", Json::string(p.recon.output));
    r.set("proj_output", Json::string(p.recon.proj_output));
    r.set("method", Json::string(method_name(p.recon.method)));
    r.set("it", Json::number(double(p.recon.it)));
    r.set("sit", Json::number(double(p.recon.sit)));
    r.set("mlog", Json::boolean(p.recon.mlog));
    r.set("weight", Json::number(p.recon.weight));
    r.set("weight_factor", Json::number(p.recon.weight_factor));
    r.set("normalize", Json::boolean(p.recon.normalize));
    r.set("simulation", Json::boolean(p.recon.simulation));
    root.set("recon", r);
    Json g = Json::object();
    g.set("so", Json::number(p.geometry.so));
    g.set("sd", Json::number(p.geometry.sd));
    g.set("angle", Json::number(p.geometry.angle));
    g.set("heli_offset", Json::number(p.geometry.heli_offset));
    g.set("heli_step", Json::number(p.geometry.heli_step));
    Json al = Json::array();
    for (float a : p.geometry.angle_list) al.arr.push_back(Json::number(a));
    g.set("angle_list", al);
    root.set("geometry", g);
    Json d = Json::object();
    d.set("px", Json::number(p.detector.px));
    d.set("rx", Json::number(p.detector.rx));
    d.set("sx", Json::number(p.detector.sx));
    d.set("sy", Json::number(p.detector.sy));
    d.set("module_number", Json::number(double(p.detector.module_number)));
    d.set("module_angle_offset", p.detector.module_angle_offset
                                     ? Json::number(*p.detector.module_angle_offset)
                                     : Json());
    d.set("module_center_offset_z", Json::number(p.detector.module_center_offset_z));
    root.set("detector", d);
    Json m2 = Json::object();
    m2.set("overlap_fct", Json::number(p.method2.overlap_fct));
    root.set("method2", m2);
    Json h = Json::object();
    h.set("devx", Json::string(p.hartmann.devx));
    h.set("devy", Json::string(p.hartmann.devy));
    h.set("iterations", Json::number(p.hartmann.iterations));
    h.set("iterations_fct", Json::number(p.hartmann.iterations_fct));
    h.set("iterations_step", Json::number(double(p.hartmann.iterations_step)));
    h.set("method", Json::string(p.hartmann.method));
    root.set("hartmann", h);
    Json i = Json::object();
    i.set("iterations", Json::number(p.ipr.iterations));
    i.set("iterations_fct", Json::number(p.ipr.iterations_fct));
    i.set("iterations_step", Json::number(double(p.ipr.iterations_step)));
    i.set("energy_kev", Json::number(p.ipr.energy_kev));
    root.set("ipr", i);
    Json x = Json::object();
    static const char *st[] = {"f32", "f16", "bf16"};
    x.set("storage", Json::string(st[p.xpr.storage]));
    x.set("final_fp32", Json::boolean(p.xpr.final_fp32));
    x.set("accumulator", p.xpr.accumulator == XPR_ACC_F32 ? Json::string("f32")
                                                          : Json::number(p.xpr.accumulator));
    x.set("interp", Json::string(p.xpr.interp == XPR_INTERP_BILINEAR ? "bilinear" : "nearest"));
    x.set("jitter", Json::boolean(p.xpr.jitter));
    x.set("jitter_seed", Json::number(double(p.xpr.jitter_seed)));
    x.set("geometry_fp64", Json::boolean(p.xpr.geometry_fp64));
    x.set("slab_height", Json::number(double(p.xpr.slab_height)));
    x.set("device_budget_gb", Json::number(p.xpr.device_budget_gb));
    x.set("layer_store", Json::string(p.xpr.layer_store));
    x.set("threads", Json::number(p.xpr.threads));
    x.set("run_log", Json::string(p.xpr.run_log));
    x.set("autotune", Json::boolean(p.xpr.autotune));
    x.set("narrow_sums", Json::boolean(p.xpr.narrow_sums));
    x.set("quiet", Json::boolean(p.xpr.quiet));
    x.set("device", Json::number(p.xpr.device));
    root.set("xpr", x);
    return root;
}

Parameters load_parameters(const std::string &arg)
{
    std::string text;
    if (!arg.empty() && arg[0] == '{') {          // parameters.hpp:59-76
        text = arg;
    } else {
        std::ifstream f(arg, std::ios::binary);
        std::stringstream ss;
        ss << f.rdbuf();
        text = ss.str();
        if (text.empty()) throw std::runtime_error("Can not read the parameters file");
    }
    return parameters_from_json(json_parse(text));
}

} // namespace xpr
