"""Runner for seispie INI configs on the MI355X solver.

    python -m seispie_amd.sprun --workdir examples/forward [--config_file config.ini]

Same contract as the reference's scripts/sprun:10-89: INI files with an
`inherit = <sibling dir>` key in [workflow] merge the parent's sections
(parent [path] entries are rewritten to ../<parent>/<value>, 'no' values
are dropped), then the workflow named by [workflow] mode is created and,
for each section in workflow.modules, the plugin named by <section>.method
is loaded from seispie_amd.<section>.<method>, given .config/.path/.mpi,
set up and attached to the workflow, which is run.

[workflow] mpi = yes is accepted and ignored: every shot runs in this
process on one GPU (the reference's MPI path needs one rank per shot).
"""
from argparse import ArgumentParser
from configparser import ConfigParser
from importlib import import_module
from os import getcwd, makedirs, path

SECTIONS = ['workflow', 'slurm', 'preprocess', 'solver', 'postprocess', 'optimize', 'path']


def parse_config(workdir, config_file):
    """scripts/sprun:10-42"""
    config_path = path.join(workdir, config_file)
    if not path.exists(workdir):
        raise FileNotFoundError(workdir)
    if not path.exists(config_path):
        raise FileNotFoundError(config_path)
    config = ConfigParser()
    config.read(config_path)
    if 'inherit' in config['workflow']:
        parent_name = config['workflow']['inherit']
        parent, _ = parse_config(workdir + '/../' + parent_name, config_file)
        for section in SECTIONS:
            if section not in parent:
                continue
            if section not in config:
                config[section] = parent[section]
                if section == 'path':
                    for key in config[section]:
                        if config[section][key] != 'no':
                            config[section][key] = '../' + parent_name + '/' + config[section][key]
            else:
                for key in parent[section]:
                    if key in config[section]:
                        continue
                    if section == 'path' and parent[section][key] != 'no':
                        config[section][key] = '../' + parent_name + '/' + parent[section][key]
                    else:
                        config[section][key] = parent[section][key]
    return config, workdir


def setup(config, workdir, run=True):
    """scripts/sprun:55-89; returns the workflow object."""
    path_dict = {}
    for key in config['path']:
        if config['path'][key] != 'no':
            path_dict[key] = path.join(workdir, config['path'][key])
    if not path.exists(path_dict['output']):
        makedirs(path_dict['output'])

    mode = config['workflow']['mode']
    workflow = getattr(import_module('seispie_amd.workflow.' + mode), mode)()
    workflow.dict = path_dict
    workflow.mpi = None
    workflow.setup()

    for section in workflow.modules:
        method = config[section]['method'] if section in config and 'method' in config[section] else 'default'
        module = getattr(import_module('seispie_amd.' + section + '.' + method), method)()
        module.config = config[section] if section in config else {}
        module.path = path_dict
        module.mpi = None
        module.setup(workflow)
        setattr(workflow, section, module)

    if run:
        workflow.run()
    return workflow


def main(argv=None):
    parser = ArgumentParser()
    parser.add_argument('--workdir', nargs='?', default=getcwd())
    parser.add_argument('--config_file', nargs='?', default='config.ini')
    parser.add_argument('--add_path', nargs='?', default='no')
    args = parser.parse_args(argv)
    config, workdir = parse_config(args.workdir, args.config_file)
    return setup(config, workdir)


if __name__ == '__main__':
    main()
